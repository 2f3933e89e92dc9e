// Oracle transport: scalar, one path at a time, restating the reference's
// hero-wavelength path tracer (shaders/packet_pt_event_body.inc,
// shaders/hero_pt_helpers.h, src/hero_spectrum_shared.h) with PBRT v4 BSDFs
// (shaders/conductor.h, shaders/dielectric.h), homogeneous media
// (shaders/medium.h) and the PBRT watertight triangle test.
// TEST INFRASTRUCTURE ONLY (see oracle_internal.h).
#include "oracle_internal.h"

#include <atomic>
#include <functional>
#include <numeric>
#include <thread>

namespace orc {

// =========================================================== scene + BVH
struct Light {
    uint32_t kind;   // 0 triangle, 1 spot, 2 environment, 3 quad (rectangle of two triangles)
    uint32_t index;  // triangle id (a quad's first triangle) or spot index
    uint32_t index2; // a quad's second triangle
    uint32_t leaf;   // node of the light BVH holding this light (RP_NONE: infinite or powerless)
    float pmf;       // power sampler: selection probability
    float threshold; // power sampler: alias threshold
    uint32_t alias;  // power sampler: alias light
    float inv_area;  // triangle / quad
    V3 p00, ex, ey;  // quad: p00 + u ex + v ey, u, v in [0, 1]
    bool rect;       // quad: passes the binary32 rectangle test (spherical sampling allowed)
    V3 ux, uy, uz;   // quad: unit edge directions and their cross product (binary32)
    float lx, ly;    // quad: edge lengths
};
// The quad's unit frame and edge lengths, as the host precomputes them.
static void quad_axes(Light &L) {
    L.lx = std::sqrt(dot(L.ex, L.ex));
    L.ly = std::sqrt(dot(L.ey, L.ey));
    L.ux = L.ex * (1.0f / L.lx);
    L.uy = L.ey * (1.0f / L.ly);
    L.uz = cross3(L.ux, L.uy);
}
// Light-BVH node (PBRT v4 LightBounds + tree links): bounds, direction cone,
// emission cutoff, power normalised by the tree's total; flags bit 0 = two-sided.
struct LightNode {
    float low[3], high[3], dir[3];
    float cos_outer, cos_emission, power;
    uint32_t child, parent, light, flags;
};

// A light's input to the tables (LightSamplingInput, src/light_sampling.hpp:11-17).
struct LightInput {
    double power = 0, bounds_power = -1;
    float low[3] = {0, 0, 0}, high[3] = {0, 0, 0}, dir[3] = {0, 0, 1};
    float cos_outer = -1, cos_emission = 0;
    bool two_sided = true, infinite = false;
};

struct Scene {
    rp_scene_desc d{};
    std::vector<float> positions;
    std::vector<uint32_t> indices, tri_object;
    std::vector<rp_object> objects;
    std::vector<rp_material> materials;
    std::vector<rp_emission> emissions;
    std::vector<rp_medium> media;
    std::vector<rp_spot_light> spots;
    std::vector<float> spectra;
    // binary BVH
    struct Node {
        float lo[3], hi[3];
        uint32_t left_or_first, count; // count == 0: interior (children left, left+1)
    };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    std::vector<Light> lights;
    std::vector<LightNode> lnodes;   // light BVH (node i + 1 is the first child of an interior node i)
    std::vector<LightInput> light_inputs; // per light: powers, bounds and cones the tables were built from
    uint32_t env_light = RP_NONE;    // the environment's light id
    float env_pmf = 0;               // its selection probability (context independent)
    std::vector<uint32_t> tri_light; // triangle -> light index or RP_NONE
    float scene_radius = 0;
    uint32_t m = 0;
    // Test switch (or_scene_set_sampling): the scene whose materials and
    // media drive every sampling decision and density -- BSDF and
    // free-flight sampling, hero-lane pdfs, MIS weights -- while this scene's
    // values are evaluated (BSDF values, scattering coefficients,
    // transmittance).  Null: this scene for both.  With it, central
    // differences of the forward render in the parameters keep every path
    // fixed (the detached estimator the adjoint differentiates).
    const Scene *sampling = nullptr;
    V3 vert(uint32_t i) const { return {positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]}; }
    void tri(uint32_t t, V3 &a, V3 &b, V3 &c) const {
        a = vert(indices[3 * t]);
        b = vert(indices[3 * t + 1]);
        c = vert(indices[3 * t + 2]);
    }
};

static void build_bvh(Scene &s) {
    uint32_t n = s.d.triangle_count;
    s.order.resize(n);
    std::iota(s.order.begin(), s.order.end(), 0u);
    std::vector<V3> cen(n), lo(n), hi(n);
    for (uint32_t t = 0; t < n; ++t) {
        V3 a, b, c;
        s.tri(t, a, b, c);
        lo[t] = {std::min({a.x, b.x, c.x}), std::min({a.y, b.y, c.y}), std::min({a.z, b.z, c.z})};
        hi[t] = {std::max({a.x, b.x, c.x}), std::max({a.y, b.y, c.y}), std::max({a.z, b.z, c.z})};
        cen[t] = {0.5f * (lo[t].x + hi[t].x), 0.5f * (lo[t].y + hi[t].y), 0.5f * (lo[t].z + hi[t].z)};
    }
    s.nodes.clear();
    s.nodes.reserve(2 * n / 2 + 2);
    struct Item { uint32_t node, begin, end; };
    std::vector<Item> stack;
    s.nodes.push_back({});
    stack.push_back({0, 0, n});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        Scene::Node nd{};
        float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = it.begin; i < it.end; ++i) {
            uint32_t t = s.order[i];
            for (int a = 0; a < 3; ++a) {
                l[a] = std::min(l[a], get(lo[t], a));
                h[a] = std::max(h[a], get(hi[t], a));
                cl[a] = std::min(cl[a], get(cen[t], a));
                ch[a] = std::max(ch[a], get(cen[t], a));
            }
        }
        // Pad boxes so that the triangle test alone decides hits.
        for (int a = 0; a < 3; ++a) {
            float pad = std::max(std::fabs(l[a]), std::fabs(h[a])) * 0x1p-18f + 1e-30f;
            nd.lo[a] = l[a] - pad;
            nd.hi[a] = h[a] + pad;
        }
        uint32_t count = it.end - it.begin;
        int axis = 0;
        float ext = ch[0] - cl[0];
        for (int a = 1; a < 3; ++a)
            if (ch[a] - cl[a] > ext) { ext = ch[a] - cl[a]; axis = a; }
        if (count <= 4 || !(ext > 0)) {
            nd.left_or_first = it.begin;
            nd.count = count;
            s.nodes[it.node] = nd;
            continue;
        }
        uint32_t mid = it.begin + count / 2;
        std::nth_element(s.order.begin() + it.begin, s.order.begin() + mid, s.order.begin() + it.end,
                         [&](uint32_t x, uint32_t y) {
                             float cx = get(cen[x], axis), cy = get(cen[y], axis);
                             return cx < cy || (cx == cy && x < y);
                         });
        uint32_t left = uint32_t(s.nodes.size());
        s.nodes.push_back({});
        s.nodes.push_back({});
        nd.left_or_first = left;
        nd.count = 0;
        s.nodes[it.node] = nd;
        stack.push_back({left, it.begin, mid});
        stack.push_back({left + 1, mid, it.end});
    }
}

// PBRT v4 watertight ray/triangle test (Woop, Benthin, Wald 2013) with the
// conservative t > 0 error bound.
static bool intersect_triangle(V3 o, V3 d, float tmax, V3 p0, V3 p1, V3 p2, float &t_out, float &b0o,
                               float &b1o, float &b2o) {
    V3 p0t = p0 - o, p1t = p1 - o, p2t = p2 - o;
    V3 ad = vabs(d);
    int kz = ad.x > ad.y ? (ad.x > ad.z ? 0 : 2) : (ad.y > ad.z ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    V3 dp = {get(d, kx), get(d, ky), get(d, kz)};
    p0t = {get(p0t, kx), get(p0t, ky), get(p0t, kz)};
    p1t = {get(p1t, kx), get(p1t, ky), get(p1t, kz)};
    p2t = {get(p2t, kx), get(p2t, ky), get(p2t, kz)};
    float sx = -dp.x / dp.z, sy = -dp.y / dp.z, sz = 1.0f / dp.z;
    p0t.x = p0t.x + sx * p0t.z;
    p0t.y = p0t.y + sy * p0t.z;
    p1t.x = p1t.x + sx * p1t.z;
    p1t.y = p1t.y + sy * p1t.z;
    p2t.x = p2t.x + sx * p2t.z;
    p2t.y = p2t.y + sy * p2t.z;
    float e0 = dop(p1t.x, p2t.y, p1t.y, p2t.x);
    float e1 = dop(p2t.x, p0t.y, p2t.y, p0t.x);
    float e2 = dop(p0t.x, p1t.y, p0t.y, p1t.x);
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        e0 = float(double(p2t.y) * double(p1t.x) - double(p2t.x) * double(p1t.y));
        e1 = float(double(p0t.y) * double(p2t.x) - double(p0t.x) * double(p2t.y));
        e2 = float(double(p1t.y) * double(p0t.x) - double(p1t.x) * double(p0t.y));
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float det = e0 + e1 + e2;
    if (det == 0.0f) return false;
    p0t.z = p0t.z * sz;
    p1t.z = p1t.z * sz;
    p2t.z = p2t.z * sz;
    float ts = e0 * p0t.z + e1 * p1t.z + e2 * p2t.z;
    if (det < 0 && (ts >= 0 || ts < tmax * det)) return false;
    if (det > 0 && (ts <= 0 || ts > tmax * det)) return false;
    float inv = 1.0f / det;
    float b0 = e0 * inv, b1 = e1 * inv, b2 = e2 * inv;
    float t = ts * inv;
    float max_zt = std::max(std::max(std::fabs(p0t.z), std::fabs(p1t.z)), std::fabs(p2t.z));
    float delta_z = gamma_n(3) * max_zt;
    float max_xt = std::max(std::max(std::fabs(p0t.x), std::fabs(p1t.x)), std::fabs(p2t.x));
    float max_yt = std::max(std::max(std::fabs(p0t.y), std::fabs(p1t.y)), std::fabs(p2t.y));
    float delta_x = gamma_n(5) * (max_xt + max_zt);
    float delta_y = gamma_n(5) * (max_yt + max_zt);
    float delta_e = 2.0f * (gamma_n(2) * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
    float max_e = std::max(std::max(std::fabs(e0), std::fabs(e1)), std::fabs(e2));
    float delta_t = 3.0f * (gamma_n(3) * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * std::fabs(inv);
    if (t <= delta_t) return false;
    t_out = t;
    b0o = b0;
    b1o = b1;
    b2o = b2;
    return true;
}

struct Hit {
    uint32_t tri = RP_NONE;
    float t = INFINITY, b0 = 0, b1 = 0, b2 = 0;
};

static bool box_hit(const Scene::Node &n, V3 o, V3 inv, float tmax) {
    float t0 = 0.0f, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        float oa = get(o, a), ia = get(inv, a);
        float tn = (n.lo[a] - oa) * ia, tf = (n.hi[a] - oa) * ia;
        if (tn > tf) std::swap(tn, tf);
        tf *= 1.0f + 2.0f * gamma_n(3);
        if (tn == tn) t0 = tn > t0 ? tn : t0;
        if (tf == tf) t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    return true;
}

// Closest hit: smallest t, ties to the smallest triangle id (order independent).
static Hit closest(const Scene &s, V3 o, V3 d, float tmax) {
    Hit h;
    if (s.nodes.empty()) return h;
    V3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const auto &n = s.nodes[stack[--sp]];
        // Cull beyond min(tmax, h.t) grown by 2^-16: a box's rounded entry
        // distance can exceed the rounded distance of a triangle inside it
        // that ties the closest hit (a shared edge), which the smallest-id
        // rule must still reach.
        if (!box_hit(n, o, inv, std::min(tmax, h.t) * (1.0f + 0x1p-16f))) continue;
        if (n.count) {
            for (uint32_t i = 0; i < n.count; ++i) {
                uint32_t tri = s.order[n.left_or_first + i];
                V3 a, b, c;
                s.tri(tri, a, b, c);
                float t, b0, b1, b2;
                if (intersect_triangle(o, d, tmax, a, b, c, t, b0, b1, b2) &&
                    (t < h.t || (t == h.t && tri < h.tri))) {
                    h = {tri, t, b0, b1, b2};
                }
            }
        } else {
            stack[sp++] = n.left_or_first;
            stack[sp++] = n.left_or_first + 1;
        }
    }
    return h;
}

// ====================================================== surface geometry
struct SurfacePoint {
    V3 p, n, err;
};
static SurfacePoint surface_point(const Scene &s, const Hit &h) {
    V3 a, b, c;
    s.tri(h.tri, a, b, c);
    SurfacePoint sp;
    sp.p = h.b0 * a + h.b1 * b + h.b2 * c;
    V3 abs_sum = vabs(h.b0 * a) + vabs(h.b1 * b) + vabs(h.b2 * c);
    sp.err = gamma_n(7) * abs_sum;
    sp.n = normalize(cross(a - c, b - c));
    return sp;
}
// PBRT v4 OffsetRayOrigin.
static V3 offset_origin(V3 p, V3 err, V3 n, V3 w) {
    float dd = dot(vabs(n), err);
    V3 off = dd * n;
    if (dot(w, n) < 0) off = -off;
    V3 po = p + off;
    float *c = &po.x;
    const float *oc = &off.x;
    for (int i = 0; i < 3; ++i) {
        if (oc[i] > 0) c[i] = next_up(c[i]);
        else if (oc[i] < 0) c[i] = next_down(c[i]);
    }
    return po;
}

struct Frame {
    V3 x, y, z;
    V3 local(V3 v) const { return {dot(v, x), dot(v, y), dot(v, z)}; }
    V3 world(V3 v) const { return x * v.x + y * v.y + z * v.z; }
};
// Duff et al. 2017 branchless orthonormal basis (PBRT v4 CoordinateSystem).
static Frame frame_from_z(V3 n) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    Frame f;
    f.x = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    f.y = {b, sign + n.y * n.y * a, -n.y};
    f.z = n;
    return f;
}

// ================================================================ BSDFs
struct Lane {
    uint32_t kind = RP_MATERIAL_DIFFUSE;
    float r = 0, eta = 1, k = 0, ax = 0, ay = 0;
    bool smooth = false;
};

static Lane prepare_lane(const Scene &s, const rp_material &m, float lambda) {
    Lane l;
    l.kind = m.kind;
    const float *pool = s.spectra.data();
    if (m.kind == RP_MATERIAL_DIFFUSE) {
        l.r = spectrum_eval(m.reflectance, lambda, pool);
        return l;
    }
    l.eta = spectrum_eval(m.eta, lambda, pool);
    if (m.kind == RP_MATERIAL_CONDUCTOR) l.k = spectrum_eval(m.k, lambda, pool);
    float u = m.roughness_u, v = m.roughness_v;
    l.ax = m.remap_roughness ? std::sqrt(u) : u;
    l.ay = m.remap_roughness ? std::sqrt(v) : v;
    if (m.kind == RP_MATERIAL_DIELECTRIC && l.eta == 0.0f) l.eta = 1.0f;
    l.smooth = std::max(l.ax, l.ay) < 1e-3f || (m.kind == RP_MATERIAL_DIELECTRIC && l.eta == 1.0f);
    if (!l.smooth) {
        l.ax = std::max(l.ax, 1e-4f);
        l.ay = std::max(l.ay, 1e-4f);
    }
    return l;
}

// ------------------------------------------------- microfacet BSDFs
// The reference's binary32 forms (shaders/conductor.h:22-208,
// shaders/dielectric.h:15-234, adapted from PBRT v4): a scaled unit vector,
// compensated Fresnel terms, the rationalised refraction coefficient and
// the compensated VNDF height.  Where the reference divides by a range scale
// (unit vectors, complex ratios, Lambda's length) this multiplies by one
// reciprocal: those quantities do not depend on the scale.  A form the
// reference marks invalid yields NaN here, which fails the path's
// arithmetic checks.
struct Sum2 {
    float s, e;
};
// Error-free sum of finite operands (Knuth's TwoSum).
static Sum2 two_sum(float a, float b) {
    const float s = a + b, bv = s - a;
    return {s, (a - (s - bv)) + (b - bv)};
}
static bool finite3(V3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
// v / |v| after scaling by the largest component (false: zero or not finite).
static bool unit3(V3 v, V3 &out) {
    const float s = std::max(std::max(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z));
    if (!(s > 0) || !std::isfinite(s)) return false;
    v = v * (1.0f / s);
    out = v * (1.0f / std::sqrt(dot(v, v)));
    return finite3(out);
}
// |a|^2 / |b|^2 of two complex numbers, scaled by b's larger component.
static bool ratio_norm(float ar, float ai, float br, float bi, float &out) {
    const float s = std::max(std::fabs(br), std::fabs(bi));
    if (!(s > 0) || !std::isfinite(s) || !std::isfinite(ar) || !std::isfinite(ai)) return false;
    const float r = 1.0f / s;
    ar = ar * r, ai = ai * r, br = br * r, bi = bi * r;
    out = (ar * ar + ai * ai) / (br * br + bi * bi);
    return std::isfinite(out);
}
// Conductor Fresnel reflectance at cos(theta) for eta + i k: the complex
// Snell root sqrt(n^2 - sin^2) with the three squares and their sum
// compensated (shaders/conductor.h:57-98).
static float fresnel_complex(float cosine, float eta, float k) {
    if (!std::isfinite(cosine) || !std::isfinite(eta) || !std::isfinite(k) || eta < 0 || k < 0 ||
        (eta == 0 && k == 0))
        return NAN;
    const float c = std::min(std::max(cosine, 0.0f), 1.0f);
    if (eta == 1 && k == 0) return 0.0f;
    if (c == 0) return 1.0f;
    const float n2x = eta * eta - k * k, n2y = 2 * eta * k;
    if (!std::isfinite(n2x) || !std::isfinite(n2y)) return NAN;
    const float eta2 = eta * eta, k2 = k * k, c2 = c * c;
    const Sum2 ab = two_sum(eta2, -k2), cd = two_sum(c2, -1.0f), total = two_sum(ab.s, cd.s);
    const float correction = ab.e + cd.e + total.e + std::fma(eta, eta, -eta2) - std::fma(k, k, -k2) +
                             std::fma(c, c, -c2);
    const float zx = total.s + correction, zy = n2y;
    const float scale = std::max(std::fabs(zx), std::fabs(zy));
    float radius = 0.0f;
    if (scale != 0) {
        const float r = 1.0f / scale, sx = zx * r, sy = zy * r;
        radius = scale * std::sqrt(sx * sx + sy * sy);
    }
    if (!std::isfinite(radius)) return NAN;
    float qr, qi;
    if (zx >= 0) {
        qr = std::sqrt(std::max(0.0f, 0.5f * radius + 0.5f * zx));
        qi = qr > 0 ? zy / (2 * qr) : 0.0f;
    } else {
        qi = std::sqrt(std::max(0.0f, 0.5f * radius - 0.5f * zx));
        qr = qi > 0 ? zy / (2 * qi) : 0.0f;
    }
    float rs, rp;
    if (!ratio_norm(c - qr, 0.0f - qi, c + qr, qi, rs) ||
        !ratio_norm(n2x * c - qr, n2y * c - qi, n2x * c + qr, n2y * c + qi, rp))
        return NAN;
    const float f = (rs + rp) * 0.5f;
    return std::isfinite(f) && f >= 0 ? f : NAN;
}
// Dielectric optics at cos(theta) (negative: from inside) for relative index
// eta (shaders/dielectric.h:29-71): reflectance F, relative index etap of the
// side the ray enters, transmitted cosine ct, total internal reflection.
// The Snell discriminant is compensated, with the authored eta kept on both
// sides.
struct Optics {
    float F, etap, ct;
    bool tir, ok;
};
static Optics optics(float cosine, float eta) {
    Optics o = {0.0f, 1.0f, 0.0f, false, false};
    if (!std::isfinite(cosine)) return o;
    o.etap = cosine < 0 ? 1 / eta : eta;
    const float c = std::min(std::fabs(cosine), 1.0f);
    if (!(o.etap > 0) || !std::isfinite(o.etap)) return o;
    if (o.etap == 1) {
        o.ct = c;
        o.ok = true;
        return o;
    }
    const float c2 = c * c, e2 = eta * eta;
    float disc;
    if (cosine < 0) {
        const Sum2 a = two_sum(1.0f, -e2);
        disc = std::fma(e2, c2, a.s) +
               (a.e + std::fma(eta, eta, -e2) * (c2 - 1) + e2 * std::fma(c, c, -c2));
    } else {
        const Sum2 a = two_sum(e2, -1.0f), b = two_sum(a.s, c2);
        disc = b.s + (a.e + b.e + std::fma(eta, eta, -e2) + std::fma(c, c, -c2));
    }
    if (!std::isfinite(disc)) return o;
    o.tir = disc <= 0;
    if (o.tir) {
        o.F = 1.0f;
        o.ok = true;
        return o;
    }
    o.ct = std::sqrt(disc) / (cosine < 0 ? 1.0f : eta);
    const float rp = (o.etap * c - o.ct) / (o.etap * c + o.ct);
    const float rs = (c - o.etap * o.ct) / (c + o.etap * o.ct);
    o.F = 0.5f * (rp * rp + rs * rs);
    o.ok = std::isfinite(o.ct) && std::isfinite(o.F) && o.F >= 0 && o.F <= 1;
    return o;
}
static float fresnel_dielectric(float cosine, float eta) {
    const Optics o = optics(cosine, eta);
    return o.ok ? o.F : NAN;
}
// c / etap - ct written with the authored index on both sides
// (shaders/dielectric.h:125-130).
static float refraction_coefficient(float eta, float co, float ct) {
    const float c = std::fabs(co);
    return co < 0 ? ((eta - 1) * (eta + 1)) / (eta * c + ct) : (((1 - eta) * (1 + eta)) / eta) / (c + eta * ct);
}
// Trowbridge-Reitz D with PBRT's grazing cutoff, and Smith Lambda.
static float ggx_d(const Lane &l, V3 m) {
    if (m.z * m.z * m.z * m.z < 1e-16f) return 0.0f;
    V3 v = {m.x / l.ax, m.y / l.ay, m.z};
    float t = dot(v, v);
    const float den = kPi * l.ax * l.ay * t * t;
    if (!std::isfinite(den) || !(den > 0)) return NAN;
    const float d = 1.0f / den;
    return d > 0 ? d : NAN;
}
static float ggx_lambda(const Lane &l, V3 w) {
    const V3 v = {1.0f, l.ax * w.x / w.z, l.ay * w.y / w.z};
    const float s = std::max(std::max(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z));
    return 0.5f * (s * length(v * (1.0f / s)) - 1.0f);
}
// Heitz 2018 visible-normal sampling; the height over the warped disk keeps
// 1 - u at the disk's edge (shaders/conductor.h:153-178).  False: invalid.
static bool ggx_sample_wm(const Lane &l, V3 w, float u0, float u1, V3 &m) {
    V3 wh;
    if (!unit3(V3{l.ax * w.x, l.ay * w.y, w.z}, wh)) return false;
    if (wh.z < 0) wh = -wh;
    V3 t1 = {1.0f, 0.0f, 0.0f};
    if (wh.z < 0.99999f && !unit3(V3{-wh.y, wh.x, 0.0f}, t1)) return false;
    const V3 t2 = cross3(wh, t1);
    const float r = std::sqrt(u0);
    float sn, cs;
    sincos_turns(u1, sn, cs);
    const float dx = r * cs, old_y = r * sn;
    const float h = std::sqrt(std::max(0.0f, 1.0f - dx * dx)), a = (1.0f + wh.z) * 0.5f;
    const float dy = (1.0f - a) * h + a * old_y;
    const float old_minus_h = old_y > 0 ? -(1.0f - u0) / (h + old_y) : old_y - h;
    const float z = std::sqrt(std::max(0.0f, (1.0f - u0) + (1.0f - a) * old_minus_h * (old_y + dy)));
    const V3 nh = dx * t1 + dy * t2 + z * wh;
    return unit3(V3{l.ax * nh.x, l.ay * nh.y, std::max(1e-6f, nh.z)}, m);
}
struct Eval {
    float f = 0, pdf = 0;
};
// Rough dielectric at (wo, wi) with its half vector (shaders/dielectric.h:72-124).
static Eval dielectric_eval(const Lane &l, V3 wo, V3 wi) {
    const float nan = NAN;
    const bool reflection = (wo.z > 0) == (wi.z > 0), entering = wo.z > 0;
    const V3 h = reflection ? wo + wi
                            : (entering ? V3{std::fma(l.eta, wi.x, wo.x), std::fma(l.eta, wi.y, wo.y),
                                             std::fma(l.eta, wi.z, wo.z)}
                                        : V3{std::fma(l.eta, wo.x, wi.x), std::fma(l.eta, wo.y, wi.y),
                                             std::fma(l.eta, wo.z, wi.z)});
    if (h.x == 0 && h.y == 0 && h.z == 0) return {0.0f, 0.0f};
    V3 m;
    if (!unit3(h, m)) return {nan, nan};
    if (m.z < 0) m = -m;
    const float co = dot(wo, m), ci = dot(wi, m);
    if (co * wo.z < 0 || ci * wi.z < 0) return {0.0f, 0.0f};
    const Optics o = optics(co, l.eta);
    if (!o.ok) return {nan, nan};
    const float lo = ggx_lambda(l, wo), li = ggx_lambda(l, wi), d = ggx_d(l, m);
    if (!std::isfinite(lo) || !std::isfinite(li) || !std::isfinite(d)) return {nan, nan};
    if (d == 0) return {0.0f, 0.0f};
    if (reflection)
        return {d * o.F / (4 * std::fabs(wo.z) * std::fabs(wi.z) * (1 + lo + li)), d * o.F / (4 * std::fabs(wo.z) * (1 + lo))};
    const float den = dot(h, m) / (entering ? l.eta : 1.0f);
    const float jacobian = std::fabs(ci) / (den * den);
    const float pdf = d * std::fabs(co) / (std::fabs(wo.z) * (1 + lo)) * jacobian * (1 - o.F);
    const float scale = entering ? 1 / (l.eta * l.eta) : l.eta * l.eta;
    return {d * (1 - o.F) * std::fabs(ci * co) * scale / (std::fabs(wo.z) * std::fabs(wi.z) * den * den * (1 + lo + li)), pdf};
}
static Eval eval_lane(const Lane &l, V3 wo, V3 wi) {
    Eval e = {0.0f, 0.0f};
    if (l.kind == RP_MATERIAL_DIFFUSE) {
        if (!(wo.z * wi.z > 0)) return e;
        e.f = l.r * kInvPi;
        e.pdf = std::fabs(wi.z) * kInvPi;
        return e;
    }
    if (l.smooth || wo.z == 0.0f || wi.z == 0.0f) return e;
    if (l.kind == RP_MATERIAL_CONDUCTOR) {
        if ((wo.z > 0) != (wi.z > 0)) return e;
        V3 m;
        if (!unit3(wo + wi, m)) return {NAN, NAN};
        const float F = fresnel_complex(std::fabs(dot(wo, m)), l.eta, l.k);
        const float lo = ggx_lambda(l, wo), li = ggx_lambda(l, wi), D = ggx_d(l, m);
        e.f = D * F / (4.0f * std::fabs(wo.z) * std::fabs(wi.z) * (1.0f + lo + li));
        e.pdf = D / (4.0f * std::fabs(wo.z) * (1.0f + lo));
        return e;
    }
    return dielectric_eval(l, wo, wi);
}
struct Sampled {
    V3 wi;
    float pdf = 0;
    float eta = 1; // relative eta of a transmission event (1 otherwise)
    bool delta = false, transmission = false, ok = false;
};
// Concentric-disk cosine hemisphere (PBRT v4 SampleCosineHemisphere).
static V3 cosine_hemisphere(float u0, float u1) {
    float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (ox != 0.0f || oy != 0.0f) {
        float r, turns;
        if (std::fabs(ox) > std::fabs(oy)) {
            r = ox;
            turns = 0.125f * (oy / ox);
        } else {
            r = oy;
            turns = 0.25f - 0.125f * (ox / oy);
        }
        float sn, cs;
        sincos_turns(turns, sn, cs);
        dx = r * cs;
        dy = r * sn;
    }
    float z = std::sqrt(std::max(0.0f, 1.0f - dx * dx - dy * dy));
    return {dx, dy, z};
}

static Sampled sample_lane(const Lane &l, V3 wo, float uc, float u0, float u1) {
    Sampled s;
    s.wi = V3{0.0f, 0.0f, 0.0f};
    s.pdf = 0.0f;
    s.eta = 1.0f;
    s.delta = s.transmission = s.ok = false;
    if (wo.z == 0.0f) return s;
    if (l.kind == RP_MATERIAL_DIFFUSE) {
        V3 wi = cosine_hemisphere(u0, u1);
        if (wo.z < 0) wi.z = -wi.z;
        s.wi = wi;
        s.pdf = std::fabs(wi.z) * kInvPi;
        s.ok = s.pdf > 0;
        return s;
    }
    if (l.kind == RP_MATERIAL_CONDUCTOR) {
        if (l.smooth) {
            s.wi = {-wo.x, -wo.y, wo.z};
            s.pdf = 1.0f;
            s.delta = true;
            s.ok = true;
            return s;
        }
        V3 m;
        if (!ggx_sample_wm(l, wo, u0, u1, m)) return s;
        const V3 wi = -wo + 2.0f * dot(wo, m) * m;
        if (!finite3(wi) || wi.z == 0.0f || (wo.z > 0) != (wi.z > 0)) return s;
        const float lo = ggx_lambda(l, wo), D = ggx_d(l, m);
        if (!(D > 0)) return s;
        s.wi = wi;
        s.pdf = D / (4.0f * std::fabs(wo.z) * (1.0f + lo));
        s.ok = std::isfinite(s.pdf) && s.pdf > 0;
        return s;
    }
    // dielectric (shaders/dielectric.h:131-228): the lobe from the Fresnel
    // reflectance at the sampled microfacet normal (the geometric normal when
    // smooth)
    V3 m = {0.0f, 0.0f, 1.0f};
    if (!l.smooth && !ggx_sample_wm(l, wo, u0, u1, m)) return s;
    const float co = dot(wo, m);
    const Optics o = optics(co, l.eta);
    if (!o.ok) return s;
    const bool reflection = uc < o.F;
    V3 wi;
    if (reflection) wi = -wo + 2.0f * co * m;
    else {
        if (o.tir) return s;
        const float coefficient = l.eta == 1 ? 0.0f : refraction_coefficient(l.eta, co, o.ct);
        wi = (co < 0 ? (-l.eta) * wo : (-wo) / l.eta) + coefficient * (co < 0 ? -m : m);
    }
    if (!finite3(wi) || wi.z == 0.0f || ((wi.z > 0) == (wo.z > 0)) != reflection) return s;
    float pdf;
    if (l.smooth) pdf = reflection ? o.F : 1 - o.F;
    else {
        const float lo = ggx_lambda(l, wo), li = ggx_lambda(l, wi), D = ggx_d(l, m);
        if (!std::isfinite(lo) || !std::isfinite(li) || !std::isfinite(D) || D == 0) return s;
        if (reflection) pdf = D * o.F / (4 * std::fabs(wo.z) * (1 + lo));
        else {
            const float den = refraction_coefficient(l.eta, co, o.ct);
            pdf = D * std::fabs(co) * o.ct * (1 - o.F) / (std::fabs(wo.z) * (1 + lo) * den * den);
        }
    }
    s.wi = wi;
    s.pdf = pdf;
    s.delta = l.smooth;
    s.transmission = !reflection;
    s.eta = reflection ? 1.0f : o.etap;
    s.ok = std::isfinite(pdf) && pdf > 0;
    return s;
}
// A companion lane's mass and density for the hero's Dirac direction wi: the
// lane keeps mass only if its own specular direction is bit for bit wi.
static void dirac_lane(const Lane &l, V3 wo, V3 wi, float uc, float u0, float u1, float &mass, float &pdf) {
    (void)uc;
    mass = pdf = 0.0f;
    if (wo.z == 0.0f || wi.z == 0.0f) return;
    bool transmission = wo.z * wi.z < 0;
    if (l.kind == RP_MATERIAL_CONDUCTOR && l.smooth) {
        if (transmission || !(wi.x == -wo.x && wi.y == -wo.y && wi.z == wo.z)) return;
        pdf = 1.0f;
        mass = fresnel_complex(std::fabs(wo.z), l.eta, l.k);
        return;
    }
    if (l.kind == RP_MATERIAL_DIELECTRIC && l.smooth) {
        // the lobe forced by the lobe uniform: F selects transmission, 0 reflection
        const float F = fresnel_dielectric(wo.z, l.eta);
        Sampled c = sample_lane(l, wo, transmission ? F : 0.0f, u0, u1);
        if (!c.ok || !(c.wi.x == wi.x && c.wi.y == wi.y && c.wi.z == wi.z)) return;
        pdf = c.pdf;
        float scale = transmission ? 1.0f / (c.eta * c.eta) : 1.0f;
        mass = (transmission ? 1 - F : F) * scale;
    }
}

// ========================================================= planted slips
// Test switch (or_plant_slip): each value plants one named estimator mistake,
// so that tests can show that their known-answer gates detect it.  Zero (the
// default) plants nothing.  The GPU renderer has no such switch.
enum Slip : int {
    SLIP_NONE = 0,
    SLIP_EMISSION_UNWEIGHTED = 1,    // emission reached by BSDF sampling added without its MIS weight
    SLIP_LIGHT_PMF_DROPPED = 2,      // emission-hit MIS: area-light pdf without the selection pmf
    SLIP_ROULETTE_UNCOMPENSATED = 3, // survivors not divided by the survival probability
    SLIP_PHASE_SIGN = 4,             // phase function of medium NEE evaluated at -cos
    SLIP_NEE_NO_TRANSMITTANCE = 5,   // medium NEE without the shadow ray's transmittance
    SLIP_ENV_PDF_HEMISPHERE = 6,     // environment MIS pdf 1/(2 pi) instead of 1/(4 pi)
    SLIP_SPOT_NO_FALLOFF = 7,        // spot light without its cone falloff
    SLIP_NEE_NO_COSINE = 8,          // surface NEE without the receiver cosine
    SLIP_HERO_TRANSMITTANCE = 9,     // every lane's pass-through transmittance at the hero's sigma_t
    SLIP_CONTINUATION_NO_COSINE = 10, // continuation weight f/pdf without |cos|
    SLIP_BVH_PMF_NO_NORMAL = 11,     // emission-hit MIS: light-BVH pmf without the receiver normal
    SLIP_RECT_PDF_UNWARPED = 12      // quad proposal density without its cosine warp
};
static std::atomic<int> g_slip{SLIP_NONE};
static bool slip(Slip v) { return g_slip.load(std::memory_order_relaxed) == v; }

// ================================================================ lights
// Light tables restated from the reference's preparation (src/metal_render.mm:
// 540-640 and 1560-1780: powers, bounds, cones) and build_light_sampling
// (src/light_sampling.cpp:226-292 unite / cost, 382-571 alias table and BVH),
// which follow PBRT v4's BVHLightSampler and PowerLightSampler.

// Power proxy of a spectrum (light_spectrum_power, src/light_sampling.cpp:17-27):
// the visible-wavelength estimate at u = 1/2 of its integral over nanometres,
// with the four wavelengths of the packet sampler (float shifts by 1/4, a
// shifted value above 1 wrapped; src/packet_spectrum_shared.h:27-58) and the
// host's binary32 atanh and cosh; zero: the plain sum over 360..830 nm.  Host
// data, not path arithmetic, so libm is used as the reference's host does.
static double spectrum_power_pool(const rp_spectrum &sp, const float *pool) {
    double v = 0;
    for (int k = 0; k < 4; ++k) {
        float up = 0.5f + float(k) * 0.25f;
        if (up > 1) up -= 1;
        float l = 538 - 138.888889f * std::atanh(0.85691062f - 1.82750197f * up);
        l = std::min(std::max(l, 360.0f), 830.0f);
        const float c = std::cosh(0.0072f * (l - 538));
        const float pdf = 0.0039398042f / (c * c);
        v += double(spectrum_eval(sp, std::round(l), pool)) / pdf / 4;
    }
    if (v == 0)
        for (int i = 0; i < 471; ++i) v += double(spectrum_eval(sp, float(360 + i), pool));
    return v;
}
static double spectrum_power(const Scene &s, const rp_spectrum &sp) { return spectrum_power_pool(sp, s.spectra.data()); }
static double spectrum_max(const Scene &s, const rp_spectrum &sp) {
    double m = 0;
    for (int i = 0; i < 471; ++i) m = std::max(m, double(spectrum_eval(sp, float(360 + i), s.spectra.data())));
    return m;
}

using D3 = std::array<double, 3>;
static D3 dsub(D3 a, D3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
static double ddot(D3 a, D3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double dlen(D3 a) { return std::sqrt(ddot(a, a)); }
static D3 dcross(D3 a, D3 b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
static D3 dvec(V3 v) { return {double(v.x), double(v.y), double(v.z)}; }

// Union of two bounds: boxes, powers, two-sidedness, the smaller emission
// cosine, and the smallest cone holding both direction cones.
static LightInput bounds_union(const LightInput &a, const LightInput &b) {
    if (a.power == 0) return b;
    if (b.power == 0) return a;
    LightInput r = a;
    r.power = a.power + b.power;
    r.two_sided = a.two_sided || b.two_sided;
    r.cos_emission = std::min(a.cos_emission, b.cos_emission);
    for (int i = 0; i < 3; ++i) {
        r.low[i] = std::min(a.low[i], b.low[i]);
        r.high[i] = std::max(a.high[i], b.high[i]);
    }
    const double pi = 3.14159265358979323846;
    const double ta = std::acos(std::clamp(double(a.cos_outer), -1.0, 1.0));
    const double tb = std::acos(std::clamp(double(b.cos_outer), -1.0, 1.0));
    const D3 wa = {a.dir[0], a.dir[1], a.dir[2]}, wb = {b.dir[0], b.dir[1], b.dir[2]};
    const double between = std::acos(std::clamp(ddot(wa, wb), -1.0, 1.0));
    if (std::min(between + tb, pi) <= ta) return r; // b's cone inside a's
    if (std::min(between + ta, pi) <= tb) {         // a's cone inside b's
        for (int i = 0; i < 3; ++i) r.dir[i] = b.dir[i];
        r.cos_outer = b.cos_outer;
        return r;
    }
    const double theta = (ta + tb + between) / 2 + 1e-6;
    if (theta >= pi) {
        r.cos_outer = -1;
        r.dir[0] = r.dir[1] = 0;
        r.dir[2] = 1;
        return r;
    }
    // rotate wa toward wb by theta - ta
    D3 t;
    for (int i = 0; i < 3; ++i) t[i] = wb[i] - ddot(wa, wb) * wa[i];
    double tl = dlen(t);
    if (tl < 1e-12) {
        int axis = std::fabs(wa[0]) < 0.8 ? 0 : 1;
        D3 e = {0, 0, 0};
        e[axis] = 1;
        for (int i = 0; i < 3; ++i) t[i] = e[i] - wa[axis] * wa[i];
        tl = dlen(t);
    }
    const double turn = theta - ta;
    for (int i = 0; i < 3; ++i) r.dir[i] = float(std::cos(turn) * wa[i] + std::sin(turn) * t[i] / tl);
    const double dl = std::sqrt(double(r.dir[0]) * r.dir[0] + double(r.dir[1]) * r.dir[1] + double(r.dir[2]) * r.dir[2]);
    for (int i = 0; i < 3; ++i) r.dir[i] = float(r.dir[i] / dl);
    r.cos_outer = std::nextafter(float(std::cos(theta)), -INFINITY);
    return r;
}

// Split cost of a cluster along `axis` of the parent: power x the cone's
// solid-angle measure x surface area, with the parent's aspect ratio.
static double bounds_cost(const LightInput &b, const LightInput &all, int axis) {
    if (b.power == 0) return 0;
    const double pi = 3.14159265358979323846;
    const double o = std::acos(std::clamp(double(b.cos_outer), -1.0, 1.0));
    const double e = std::acos(std::clamp(double(b.cos_emission), -1.0, 1.0));
    const double w = std::min(o + e, pi);
    const double omega = 2 * pi * (1 - b.cos_outer) +
                         pi / 2 * (2 * w * std::sin(o) - std::cos(o - 2 * w) - 2 * o * std::sin(o) + b.cos_outer);
    double d[3], ad[3];
    for (int i = 0; i < 3; ++i) {
        d[i] = double(b.high[i]) - double(b.low[i]);
        ad[i] = double(all.high[i]) - double(all.low[i]);
    }
    if (ad[axis] <= 0) return INFINITY;
    const double longest = std::max(ad[0], std::max(ad[1], ad[2]));
    return b.power * omega * (longest / ad[axis]) * 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

// The reference's rectangle classification of a patch p = (p00, p10, p01, p11)
// (src/light_sampling.cpp:294-311): planar, equal corner distances from the centre.
static bool patch_rectangle(const D3 p[4]) {
    const D3 n = dcross(dsub(p[1], p[0]), dsub(p[2], p[0]));
    const double nl = dlen(n), diag = dlen(dsub(p[3], p[0]));
    bool ok = nl > 0 && diag > 0 && std::fabs(ddot(dsub(p[3], p[0]), n) / (diag * nl)) <= 1e-5;
    D3 c = {0, 0, 0};
    for (int i = 0; i < 4; ++i)
        for (int a = 0; a < 3; ++a) c[a] += p[i][a] / 4;
    const double d0 = ddot(dsub(p[0], c), dsub(p[0], c));
    const int next[4] = {2, 0, 3, 1};
    for (int i = 0; i < 4; ++i) {
        const D3 dl = dsub(p[i], c);
        ok = ok && dlen(dsub(p[i], p[next[i]])) > 0 && d0 > 0 && std::fabs(ddot(dl, dl) - d0) / d0 <= 1e-4;
    }
    return ok;
}
// The binary32 affine-rectangle test of the sampler (an exact corner and a right angle).
static bool rect_f(V3 p00, V3 p10, V3 p01, V3 p11) {
    const V3 ex = p10 - p00, ey = p01 - p00;
    const float x2 = dot(ex, ex), y2 = dot(ey, ey);
    if (!(x2 > 0 && y2 > 0)) return false;
    const V3 e = p11 - (p00 + ex + ey);
    return dot(e, e) <= 0x1p-40f * std::max(x2, y2) && std::fabs(dot(ex, ey)) <= 0x1p-20f * std::sqrt(x2 * y2);
}

// The smallest-angle cone about the mean unit normal holding n normals
// (light_normal_cone, src/light_sampling.cpp:356-381): a degenerate normal or
// a vanishing mean leaves the full sphere (cos_outer = -1).
static void normal_cone(LightInput &b, const D3 *ns, int n) {
    D3 axis = {0, 0, 0};
    for (int k = 0; k < n; ++k) {
        const double l = dlen(ns[k]);
        if (!(l > 0 && std::isfinite(l))) {
            b.cos_outer = -1;
            return;
        }
        for (int a = 0; a < 3; ++a) axis[a] += ns[k][a] / l;
    }
    const double al = dlen(axis);
    if (!(al > 0)) {
        b.cos_outer = -1;
        return;
    }
    for (int a = 0; a < 3; ++a) axis[a] /= al;
    double cmin = 1;
    for (int k = 0; k < n; ++k) cmin = std::min(cmin, ddot(axis, ns[k]) / dlen(ns[k]));
    for (int a = 0; a < 3; ++a) b.dir[a] = float(axis[a]);
    b.cos_outer = std::max(-1.0f, std::nextafter(float(std::clamp(cmin, -1.0, 1.0)), -INFINITY));
}

static void build_lights(Scene &s) {
    const double pi = 3.14159265358979323846;
    s.lights.clear();
    s.lnodes.clear();
    s.light_inputs.clear();
    s.env_light = RP_NONE;
    s.tri_light.assign(s.d.triangle_count, RP_NONE);
    std::vector<LightInput> in;
    auto box_of = [](LightInput &b, const V3 *c, int n) {
        for (int a = 0; a < 3; ++a) {
            b.low[a] = INFINITY;
            b.high[a] = -INFINITY;
        }
        for (int k = 0; k < n; ++k)
            for (int a = 0; a < 3; ++a) {
                b.low[a] = std::min(b.low[a], get(c[k], a));
                b.high[a] = std::max(b.high[a], get(c[k], a));
            }
    };
    // area lights: emissive triangles in id order; (a,b,c) followed by (a,c,d)
    // of the same object forming a rectangle are one quad light
    for (uint32_t t = 0; t < s.d.triangle_count; ++t) {
        const uint32_t o = s.tri_object[t];
        const auto &obj = s.objects[o];
        if (obj.emission == RP_NONE) continue;
        const auto &em = s.emissions[obj.emission];
        const double mean = spectrum_power(s, em.radiance), bound = spectrum_max(s, em.radiance);
        const uint32_t *ia = &s.indices[3 * size_t(t)];
        if (t + 1 < s.d.triangle_count && s.tri_object[t + 1] == o && s.indices[3 * size_t(t) + 3] == ia[0] &&
            s.indices[3 * size_t(t) + 4] == ia[2]) {
            const V3 c[4] = {s.vert(ia[0]), s.vert(ia[1]), s.vert(s.indices[3 * size_t(t) + 5]), s.vert(ia[2])};
            const D3 p[4] = {dvec(c[0]), dvec(c[1]), dvec(c[2]), dvec(c[3])};
            if (patch_rectangle(p) && rect_f(c[0], c[1], c[2], c[3])) {
                const double area = dlen(dsub(p[2], p[0])) * dlen(dsub(p[1], p[0]));
                LightInput b;
                box_of(b, c, 4);
                b.power = pi * area * mean * (em.two_sided ? 2.0 : 1.0);
                b.bounds_power = pi * area * bound;
                b.two_sided = em.two_sided != 0;
                // normal cone of the four corner normals
                const D3 ns[4] = {dcross(dsub(p[1], p[0]), dsub(p[2], p[0])), dcross(dsub(p[1], p[0]), dsub(p[3], p[1])),
                                  dcross(dsub(p[3], p[2]), dsub(p[2], p[0])), dcross(dsub(p[3], p[2]), dsub(p[3], p[1]))};
                normal_cone(b, ns, 4);
                Light L{3u, t, t + 1, RP_NONE, 0, 1, 0, float(1.0 / area), c[0], c[1] - c[0], c[2] - c[0], true,
                        {}, {}, {}, 0, 0};
                quad_axes(L);
                s.tri_light[t] = s.tri_light[t + 1] = uint32_t(s.lights.size());
                s.lights.push_back(L);
                in.push_back(b);
                ++t;
                continue;
            }
        }
        V3 c[3];
        s.tri(t, c[0], c[1], c[2]);
        const D3 n = dcross(dsub(dvec(c[1]), dvec(c[0])), dsub(dvec(c[2]), dvec(c[0])));
        const double area = 0.5 * dlen(n);
        if (!(area > 0)) continue;
        LightInput b;
        box_of(b, c, 3);
        b.power = pi * area * mean * (em.two_sided ? 2.0 : 1.0);
        b.bounds_power = pi * area * bound;
        b.two_sided = em.two_sided != 0;
        for (int a = 0; a < 3; ++a) b.dir[a] = float(n[a] / (2 * area));
        b.cos_outer = std::nextafter(1.0f, 0.0f);
        s.tri_light[t] = uint32_t(s.lights.size());
        s.lights.push_back({0u, t, RP_NONE, RP_NONE, 0, 1, 0, float(1.0 / area), {}, {}, {}, false, {}, {}, {}, 0, 0});
        in.push_back(b);
    }
    for (uint32_t i = 0; i < s.spots.size(); ++i) {
        const auto &sp = s.spots[i];
        LightInput b;
        for (int a = 0; a < 3; ++a) b.low[a] = b.high[a] = sp.position[a];
        b.power = 2 * pi * ((1 - double(sp.cos_falloff_start)) + (double(sp.cos_falloff_start) - sp.cos_falloff_end) / 2) *
                  spectrum_power(s, sp.intensity);
        b.bounds_power = 4 * pi * spectrum_max(s, sp.intensity);
        b.two_sided = false;
        float ce = float(std::cos(std::acos(double(sp.cos_falloff_end)) - std::acos(double(sp.cos_falloff_start))));
        if (ce == 1) ce = sp.cos_falloff_end != sp.cos_falloff_start ? 0.999f : std::nextafter(1.0f, 0.0f);
        for (int a = 0; a < 3; ++a) b.dir[a] = sp.direction[a];
        b.cos_outer = float(std::cos(std::min(pi, std::acos(std::clamp(double(sp.cos_falloff_start), -1.0, 1.0)) + 2e-6)));
        b.cos_outer = std::max(-1.0f, std::nextafter(b.cos_outer, -INFINITY));
        b.cos_emission = ce;
        s.lights.push_back({1u, i, RP_NONE, RP_NONE, 0, 1, 0, 0.0f, {}, {}, {}, false, {}, {}, {}, 0, 0});
        in.push_back(b);
    }
    if (s.d.has_environment) {
        LightInput b;
        b.infinite = true;
        const double r = s.scene_radius;
        b.power = 4 * pi * pi * r * r * spectrum_power(s, s.d.environment);
        s.env_light = uint32_t(s.lights.size());
        s.lights.push_back({2u, 0u, RP_NONE, RP_NONE, 0, 1, 0, 0.0f, {}, {}, {}, false, {}, {}, {}, 0, 0});
        in.push_back(b);
    }
    s.light_inputs = in;
    const size_t n = in.size();
    if (n == 0) return;
    // alias table over the powers (Walker / Vose, small and large stacks)
    double total = 0, bounds_total = 0;
    std::vector<double> prob(n);
    std::vector<LightInput> bv = in;
    std::vector<uint32_t> finite;
    for (size_t i = 0; i < n; ++i) {
        bv[i].power = in[i].bounds_power < 0 ? in[i].power : in[i].bounds_power;
        if (bv[i].power == 0 && in[i].power > 0) bv[i].power = in[i].power;
        prob[i] = in[i].power == 0 && bv[i].power > 0 ? bv[i].power : in[i].power;
        total += prob[i];
        if (!in[i].infinite) bounds_total += bv[i].power;
        if (!in[i].infinite && bv[i].power > 0) finite.push_back(uint32_t(i));
    }
    std::vector<uint32_t> small, large;
    for (size_t i = 0; i < n; ++i) {
        const double pr = total > 0 ? prob[i] / total : 1.0 / double(n);
        s.lights[i].pmf = float(pr);
        prob[i] = pr * double(n);
        (prob[i] < 1 ? small : large).push_back(uint32_t(i));
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t a = small.back();
        small.pop_back();
        const uint32_t b = large.back();
        large.pop_back();
        s.lights[a].threshold = float(prob[a]);
        s.lights[a].alias = b;
        prob[b] -= 1 - prob[a];
        (prob[b] < 1 ? small : large).push_back(b);
    }
    for (uint32_t i : small) s.lights[i].threshold = 1, s.lights[i].alias = i;
    for (uint32_t i : large) s.lights[i].threshold = 1, s.lights[i].alias = i;
    // the BVH over the finite lights: 12 buckets per axis, splits 1..10 scanned
    auto bucket_of = [&](uint32_t id, int axis, const double lo[3], const double hi[3]) {
        const double c = (double(bv[id].low[axis]) + bv[id].high[axis]) / 2;
        return std::min(11, int(12 * (c - lo[axis]) / (hi[axis] - lo[axis])));
    };
    std::function<uint32_t(size_t, size_t, uint32_t, unsigned)> build = [&](size_t b, size_t e, uint32_t parent,
                                                                              unsigned depth) -> uint32_t {
        const uint32_t index = uint32_t(s.lnodes.size());
        s.lnodes.emplace_back();
        LightInput all;
        double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = b; k < e; ++k) {
            const LightInput &v = bv[finite[k]];
            all = bounds_union(all, v);
            for (int a = 0; a < 3; ++a) {
                const double c = (double(v.low[a]) + v.high[a]) / 2;
                clo[a] = std::min(clo[a], c);
                chi[a] = std::max(chi[a], c);
            }
        }
        LightNode node{};
        for (int a = 0; a < 3; ++a) {
            node.low[a] = all.low[a];
            node.high[a] = all.high[a];
            node.dir[a] = all.dir[a];
        }
        node.cos_outer = all.cos_outer;
        node.cos_emission = all.cos_emission;
        node.power = float(all.power / bounds_total);
        node.parent = parent;
        node.light = RP_NONE;
        node.child = RP_NONE;
        node.flags = all.two_sided ? 1u : 0u;
        if (e - b == 1) {
            node.light = finite[b];
            s.lights[finite[b]].leaf = index;
        } else {
            double best = INFINITY;
            int best_axis = 0, best_split = 0;
            bool found = false;
            if (depth < 48)
                for (int axis = 0; axis < 3; ++axis) {
                    if (clo[axis] == chi[axis]) continue;
                    LightInput bucket[12];
                    for (size_t k = b; k < e; ++k) {
                        const uint32_t id = finite[k];
                        const int q = bucket_of(id, axis, clo, chi);
                        bucket[q] = bounds_union(bucket[q], bv[id]);
                    }
                    for (int split = 1; split < 11; ++split) {
                        LightInput left, right;
                        for (int k = 0; k <= split; ++k) left = bounds_union(left, bucket[k]);
                        for (int k = split + 1; k < 12; ++k) right = bounds_union(right, bucket[k]);
                        const double c = bounds_cost(left, all, axis) + bounds_cost(right, all, axis);
                        if (c > 0 && c < best) {
                            best = c;
                            best_axis = axis;
                            best_split = split;
                            found = true;
                        }
                    }
                }
            size_t mid = b + (e - b) / 2;
            if (found) {
                auto it = std::stable_partition(finite.begin() + ptrdiff_t(b), finite.begin() + ptrdiff_t(e),
                                                [&](uint32_t id) { return bucket_of(id, best_axis, clo, chi) <= best_split; });
                const size_t cand = size_t(it - finite.begin());
                if (cand > b && cand < e) mid = cand;
            }
            build(b, mid, index, depth + 1);
            node.child = build(mid, e, index, depth + 1);
        }
        s.lnodes[index] = node;
        return index;
    };
    if (!finite.empty()) build(0, finite.size(), RP_NONE, 0);
    if (s.env_light != RP_NONE) {
        const uint32_t sampler = s.d.light_sampler;
        s.env_pmf = sampler == RP_LIGHT_SAMPLER_UNIFORM ? 1.0f / float(n)
                    : sampler == RP_LIGHT_SAMPLER_POWER ? s.lights[s.env_light].pmf
                                                        : 1.0f / float(1 + (s.lnodes.empty() ? 0 : 1));
    }
}

static void compute_radius(Scene &s) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = 0; v < s.d.vertex_count; ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], double(s.positions[3 * v + a]));
            hi[a] = std::max(hi[a], double(s.positions[3 * v + a]));
        }
    double r2 = 0;
    for (int a = 0; a < 3; ++a) {
        double h = 0.5 * (hi[a] - lo[a]);
        r2 += h * h;
    }
    s.scene_radius = s.d.vertex_count ? float(std::sqrt(r2)) : 0.0f;
}

// ============================================================ transport
struct Ctx {
    const Scene *s;
    uint64_t seed;
    uint32_t mode;
    uint32_t emergency;
    uint32_t boundary;
};

static std::array<uint32_t, 4> draws(const Ctx &c, uint64_t index, uint32_t block) {
    uint32_t seed_bits = uint32_t(c.seed);
    uint32_t d = 4 * block;
    return {sobol_word(index, d, c.mode, seed_bits), sobol_word(index, d + 1, c.mode, seed_bits),
            sobol_word(index, d + 2, c.mode, seed_bits), sobol_word(index, d + 3, c.mode, seed_bits)};
}
// Philox medium stream, shaders/render.metal:128-136 (medium_local_block / medium_stream_v1).
static std::array<uint32_t, 4> medium_draws(const Ctx &c, uint32_t pixel, uint64_t sample, uint32_t &draw) {
    return philox({pixel, uint32_t(sample), uint32_t(sample >> 32), 0x20000000u | draw++},
                  {uint32_t(c.seed), uint32_t(c.seed >> 32) ^ 0x6d656431u});
}

static float mixture(const float d[4], const float p[4]) {
    return (d[0] * p[0] + d[1] * p[1]) + (d[2] * p[2] + d[3] * p[3]);
}
static float dsum(const float d[4]) { return (d[0] + d[1]) + (d[2] + d[3]); }
static void normalize_density(float d[4]) {
    float m = std::max(std::max(d[0], d[1]), std::max(d[2], d[3]));
    if (!(m > 0) || !std::isfinite(m)) return;
    int e;
    (void)std::frexp(m, &e);
    for (int k = 0; k < 4; ++k) d[k] = std::ldexp(d[k], -e);
}
// Cumulative balance update (src/hero_spectrum_shared.h:116-139).
static bool advance(float density[4], float beta[4], const float num[4], const float pdf[4]) {
    float nd[4];
    for (int k = 0; k < 4; ++k) nd[k] = density[k] * pdf[k];
    float before = dsum(density), after = dsum(nd);
    if (!(before > 0 && after > 0) || !std::isfinite(after)) return false;
    float ratio = before / after;
    for (int k = 0; k < 4; ++k) {
        beta[k] = (beta[k] * num[k]) * ratio;
        if (!std::isfinite(beta[k])) return false;
        density[k] = nd[k];
    }
    normalize_density(density);
    return true;
}
static bool positive_normal(float v) {
    uint32_t w = std::bit_cast<uint32_t>(v);
    return w >= 0x00800000u && w < 0x7f800000u;
}
// Throughput roulette (shaders/hero_pt_helpers.h:56-85, src/packet_roulette_shared.h:61-108).
static bool hero_continue(float beta[4], const float eta_scale[4], bool enabled, uint32_t depth, float u) {
    if (!enabled || depth <= 1) return true;
    float maximum = 0.0f;
    for (int k = 0; k < 4; ++k) {
        float v = beta[k] * eta_scale[k];
        if (!std::isfinite(v) || v < 0 || (beta[k] > 0 && !positive_normal(v))) return true;
        maximum = std::max(maximum, v);
    }
    if (!positive_normal(maximum) || maximum >= 1.0f) return true;
    float q = 1.0f - maximum;
    float survival = 1.0f - q;
    if (!(q > 0 && q < 1) || !positive_normal(survival)) return true;
    float surv[4];
    for (int k = 0; k < 4; ++k) {
        surv[k] = beta[k];
        if (std::bit_cast<uint32_t>(beta[k]) == 0) continue;
        surv[k] = slip(SLIP_ROULETTE_UNCOMPENSATED) ? beta[k] : beta[k] / survival;
        if (!positive_normal(surv[k])) return true;
    }
    if (u < q) return false;
    for (int k = 0; k < 4; ++k) beta[k] = surv[k];
    return true;
}

struct MediumCoeff {
    float st[4], ss[4];
    float g;
};
static MediumCoeff medium_coeff(const Scene &s, uint32_t m, const Wavelengths &wl) {
    MediumCoeff c{};
    const auto &md = s.media[m];
    for (int k = 0; k < 4; ++k) {
        float a = spectrum_eval(md.sigma_a, wl.lambda[k], s.spectra.data());
        float sc = spectrum_eval(md.sigma_s, wl.lambda[k], s.spectra.data());
        c.ss[k] = sc;
        c.st[k] = a + sc;
    }
    c.g = std::min(std::max(md.g, -0.99f), 0.99f);
    return c;
}
// Henyey-Greenstein, PBRT outward convention: cos = dot(wo, wi).
static float hg(float cos_t, float g) {
    float denom = 1.0f + g * g + 2.0f * g * cos_t;
    return kInv4Pi * (1.0f - g * g) / (denom * std::sqrt(denom));
}
static V3 hg_sample(V3 wo, float g, float u0, float u1, float &pdf) {
    float cos_t;
    if (std::fabs(g) < 1e-3f) cos_t = 1.0f - 2.0f * u0;
    else {
        float sq = (1.0f - g * g) / (1.0f + g - 2.0f * g * u0);
        cos_t = -1.0f / (2.0f * g) * (1.0f + g * g - sq * sq);
    }
    cos_t = std::min(std::max(cos_t, -1.0f), 1.0f);
    float sin_t = std::sqrt(std::max(0.0f, 1.0f - cos_t * cos_t));
    float sn, cs;
    sincos_turns(u1, sn, cs);
    Frame f = frame_from_z(wo);
    V3 wi = f.world({sin_t * cs, sin_t * sn, cos_t});
    pdf = hg(cos_t, g);
    return wi;
}

static uint32_t outgoing_medium(const rp_object &o, uint32_t current, V3 n, V3 w) {
    if (o.medium_inside == o.medium_outside) return current;
    return dot(n, w) > 0 ? o.medium_outside : o.medium_inside;
}

// Visibility with transmittance through null interfaces and homogeneous media.
// Returns false when blocked by a scattering surface.
static bool transmittance(const Ctx &c, V3 o, V3 d, float tmax, uint32_t medium, const Wavelengths &wl,
                          float tr[4], uint32_t &crossings, float *mlen = nullptr, uint32_t mlen_medium = RP_NONE) {
    const Scene &s = *c.s;
    for (int k = 0; k < 4; ++k) tr[k] = 1.0f;
    if (mlen) *mlen = 0.0f;
    for (;;) {
        Hit h = closest(s, o, d, tmax);
        float seg = h.tri == RP_NONE ? tmax : h.t;
        if (medium != RP_NONE) {
            MediumCoeff mc = medium_coeff(s, medium, wl);
            float len = length(d);
            for (int k = 0; k < 4; ++k) tr[k] = tr[k] * exp_f(-mc.st[k] * (seg * len));
            if (mlen && medium == mlen_medium) *mlen = *mlen + seg * len;
        }
        if (h.tri == RP_NONE) return true;
        const auto &obj = s.objects[s.tri_object[h.tri]];
        if (s.materials[obj.material].kind != RP_MATERIAL_INTERFACE) return false;
        if (++crossings > c.boundary) return false;
        SurfacePoint sp = surface_point(s, h);
        medium = outgoing_medium(obj, medium, sp.n, d);
        V3 no = offset_origin(sp.p, sp.err, sp.n, d);
        // Continue toward the same target point.
        V3 target = o + tmax * d;
        if (std::isinf(tmax)) {
            o = no;
        } else {
            d = target - no;
            o = no;
            tmax = 1.0f;
        }
    }
}

struct PathOut {
    uint32_t status = RP_PATH_SUCCESS;
    float L[4] = {0, 0, 0, 0};
    Wavelengths wl{};
    uint64_t index = 0;
    uint32_t primary = RP_NONE;
    uint32_t vertices = 0;
    uint32_t w0 = 0, w1 = 0;
    uint64_t ext_rays = 0, shadow_rays = 0;
};

struct LightSample {
    V3 p, n;
    V3 err;
    float le[4] = {0, 0, 0, 0};
    float pdf = 0; // solid angle, including selection pmf
    bool delta = false, infinite = false, ok = false;
    V3 wi;
    float dist = 0;
};

// --- light selection (shaders/light_sampling.h:19-127, PBRT BVHLightSampler)
// Importance of a light-BVH node for a receiving point p with normal n
// (n = 0: a medium vertex): power x the cosine bound of the direction cone
// against the bounding sphere, over the squared distance.
static float light_importance(const LightNode &b, V3 p, V3 n) {
    const V3 lo = {b.low[0], b.low[1], b.low[2]}, hi = {b.high[0], b.high[1], b.high[2]};
    const V3 c = (lo + hi) * 0.5f;
    const V3 delta = p - c;
    const float r2 = dot(hi - lo, hi - lo) * 0.25f, d2 = dot(delta, delta);
    const float radius = std::sqrt(r2), dist = std::sqrt(d2);
    const float cb = d2 <= r2 ? -1.0f : std::sqrt(std::max(0.0f, 1.0f - r2 / d2));
    const float sb = std::sqrt(std::max(0.0f, 1.0f - cb * cb));
    const V3 wi = dist > 0 ? delta * (1.0f / dist) : V3{0, 0, 1};
    float cw = std::clamp(dot(V3{b.dir[0], b.dir[1], b.dir[2]}, wi), -1.0f, 1.0f);
    if (b.flags & 1u) cw = std::fabs(cw);
    const float sw = std::sqrt(std::max(0.0f, 1.0f - cw * cw));
    const float co = b.cos_outer, so = std::sqrt(std::max(0.0f, 1.0f - co * co));
    const float cx = cw > co ? 1.0f : cw * co + sw * so, sx = cw > co ? 0.0f : sw * co - cw * so;
    const float cp = cx > cb ? 1.0f : cx * cb + sx * sb;
    if (cp <= b.cos_emission) return 0.0f;
    float value = b.power * cp / std::max(d2, std::max(radius, FLT_MIN));
    if (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f) {
        const float ci = std::clamp(std::fabs(dot(wi, n)), 0.0f, 1.0f);
        const float si = std::sqrt(std::max(0.0f, 1.0f - ci * ci));
        value = value * (ci > cb ? 1.0f : ci * cb + si * sb);
    }
    return std::max(0.0f, value);
}

struct Selected {
    uint32_t light = RP_NONE;
    float pmf = 0.0f;
};
static Selected light_select(const Scene &s, V3 p, V3 n, float u) {
    const uint32_t count = uint32_t(s.lights.size());
    if (count == 0) return {};
    u = std::clamp(u, 0.0f, 0x1.fffffep-1f);
    const uint32_t sampler = s.d.light_sampler;
    if (sampler == RP_LIGHT_SAMPLER_UNIFORM) return {std::min(uint32_t(u * float(count)), count - 1), 1.0f / float(count)};
    if (sampler == RP_LIGHT_SAMPLER_POWER) {
        const float v = u * float(count);
        const uint32_t i = std::min(uint32_t(v), count - 1);
        const uint32_t id = v - float(i) < s.lights[i].threshold ? i : s.lights[i].alias;
        return {id, s.lights[id].pmf};
    }
    const uint32_t inf = s.env_light != RP_NONE ? 1u : 0u, tree = s.lnodes.empty() ? 0u : 1u;
    const float infinite = float(inf) / float(inf + tree);
    if (u < infinite) return {s.env_light, 1.0f / float(inf + tree)};
    if (!tree) return {};
    u = std::min((u - infinite) / (1.0f - infinite), 0x1.fffffep-1f);
    float pmf = 1.0f - infinite;
    uint32_t i = 0;
    for (size_t steps = 0; steps < s.lnodes.size(); ++steps) {
        const LightNode &node = s.lnodes[i];
        if (node.light != RP_NONE) {
            if (light_importance(node, p, n) <= 0.0f) return {};
            return {node.light, pmf};
        }
        const float a = light_importance(s.lnodes[i + 1], p, n), b = light_importance(s.lnodes[node.child], p, n);
        const float sum = a + b;
        if (!(sum > 0.0f)) return {};
        const float q = a / sum;
        if (u < q) {
            pmf = pmf * q;
            u = std::min(u / q, 0x1.fffffep-1f);
            i = i + 1;
        } else {
            pmf = pmf * (1.0f - q);
            u = std::min((u - q) / (1.0f - q), 0x1.fffffep-1f);
            i = node.child;
        }
    }
    return {};
}
static float light_pmf(const Scene &s, V3 p, V3 n, uint32_t id) {
    const uint32_t count = uint32_t(s.lights.size());
    if (id >= count) return 0.0f;
    const uint32_t sampler = s.d.light_sampler;
    if (sampler == RP_LIGHT_SAMPLER_UNIFORM) return 1.0f / float(count);
    if (sampler == RP_LIGHT_SAMPLER_POWER) return s.lights[id].pmf;
    const uint32_t inf = s.env_light != RP_NONE ? 1u : 0u, tree = s.lnodes.empty() ? 0u : 1u;
    if (id == s.env_light) return 1.0f / float(inf + tree);
    uint32_t i = s.lights[id].leaf;
    if (i >= s.lnodes.size()) return 0.0f;
    if (light_importance(s.lnodes[i], p, n) <= 0.0f) return 0.0f;
    float pmf = 1.0f - float(inf) / float(inf + tree);
    for (size_t steps = 0; steps < s.lnodes.size(); ++steps) {
        const uint32_t parent = s.lnodes[i].parent;
        if (parent == RP_NONE) return pmf;
        const float a = light_importance(s.lnodes[parent + 1], p, n);
        const float b = light_importance(s.lnodes[s.lnodes[parent].child], p, n);
        const float sum = a + b;
        if (!(sum > 0.0f)) return 0.0f;
        pmf = pmf * (i == parent + 1 ? a / sum : b / sum);
        i = parent;
    }
    return 0.0f;
}

// --- quad emitters: spherical-rectangle proposal (Urena, Fajardo, King 2013,
// as adapted in src/light_geometry_sampling_shared.h:174-260 and
// shaders/light_geometry_sampling.h:118-209), warped by the receiver's
// corner cosines (bilinear, PBRT SampleBilinear) when a normal is given.
static V3 unit(V3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
static float angle_between(V3 a, V3 b) {
    return dot(a, b) < 0 ? kPi - 2.0f * asin_f(std::min(1.0f, length(a + b) / 2.0f))
                         : 2.0f * asin_f(std::min(1.0f, length(a - b) / 2.0f));
}
// Area of the spherical triangle of unit vectors a, b, c (Van Oosterom-Strackee).
static float spherical_triangle_area(V3 a, V3 b, V3 c) {
    return std::fabs(2.0f * atan2_f(dot(a, cross3(b, c)), 1.0f + dot(a, b) + dot(a, c) + dot(b, c)));
}
struct Bilinear {
    float a, b, c, d; // weights at (0,0), (1,0), (0,1), (1,1)
};
static Bilinear bilinear_normalized(Bilinear w) {
    const float m = std::max(std::max(w.a, w.b), std::max(w.c, w.d));
    return m > 0 ? Bilinear{w.a / m, w.b / m, w.c / m, w.d / m} : Bilinear{1, 1, 1, 1};
}
// Inverse CDF of the density proportional to (1 - x) a + x b on [0, 1].
static float linear_warp(float u, float a, float b) {
    if (u == 0.0f && a == 0.0f) return 0.0f;
    return std::min(0x1.fffffep-1f, u * (a + b) / (a + std::sqrt((1.0f - u) * a * a + u * b * b)));
}
static void bilinear_warp(float &u0, float &u1, Bilinear w) {
    w = bilinear_normalized(w);
    const float y = linear_warp(u1, w.a + w.b, w.c + w.d);
    const float x = linear_warp(u0, (1.0f - y) * w.a + y * w.c, (1.0f - y) * w.b + y * w.d);
    u0 = x;
    u1 = y;
}
static float bilinear_density(float x, float y, Bilinear w) {
    if (x < 0 || x > 1 || y < 0 || y > 1) return 0.0f;
    w = bilinear_normalized(w);
    return 4.0f * ((1.0f - x) * (1.0f - y) * w.a + x * (1.0f - y) * w.b + (1.0f - x) * y * w.c + x * y * w.d) /
           (w.a + w.b + w.c + w.d);
}
// The rectangle seen from r in its own frame (x along ex, y along ey, z
// toward r negative): extents, the spherical-rectangle constants, the solid
// angle, and the unit directions to the four corners.
struct RectFrame {
    float x0, x1, y0, y1, z0, b0, b1, g2, g3, area, zs;
    V3 w00, w10, w01, w11;
};
static RectFrame rect_frame(const Light &L, V3 r) {
    RectFrame f;
    const V3 delta = L.p00 - r;
    f.x0 = dot(delta, L.ux);
    f.x1 = f.x0 + L.lx;
    f.y0 = dot(delta, L.uy);
    f.y1 = f.y0 + L.ly;
    f.z0 = dot(delta, L.uz);
    f.zs = 1.0f;
    if (f.z0 > 0) {
        f.z0 = -f.z0;
        f.zs = -1.0f;
    }
    const V3 v00 = {f.x0, f.y0, f.z0}, v10 = {f.x1, f.y0, f.z0}, v11 = {f.x1, f.y1, f.z0}, v01 = {f.x0, f.y1, f.z0};
    const V3 n0 = unit(cross3(v00, v10)), n2 = unit(cross3(v11, v01)), n3 = unit(cross3(v01, v00));
    f.g2 = angle_between(-n2, n3);
    f.g3 = angle_between(-n3, n0);
    f.b0 = n0.z;
    f.b1 = n2.z;
    f.w00 = unit(v00);
    f.w10 = unit(v10);
    f.w01 = unit(v01);
    f.w11 = unit(v11);
    f.area = spherical_triangle_area(f.w00, f.w10, f.w11) + spherical_triangle_area(f.w00, f.w11, f.w01);
    return f;
}
// Corner cosines |n . w| (floored at 0.01) of the receiver normal n.
static Bilinear rect_cosines(const RectFrame &f, const Light &L, V3 n) {
    const V3 nl = {dot(n, L.ux), dot(n, L.uy), f.zs * dot(n, L.uz)};
    return {std::max(0.01f, std::fabs(dot(f.w00, nl))), std::max(0.01f, std::fabs(dot(f.w10, nl))),
            std::max(0.01f, std::fabs(dot(f.w01, nl))), std::max(0.01f, std::fabs(dot(f.w11, nl)))};
}
// Square sample -> rectangle coordinates (u along ex, v along ey).
static void rect_sample(const RectFrame &f, float s0, float s1, float &u, float &v) {
    const float au = s0 * f.area - f.g2 - f.g3;
    const float fu = (cos_f(au) * f.b0 - f.b1) / sin_f(au);
    const float cu = std::min(0x1.fffffep-1f, std::max(-0x1.fffffep-1f, std::copysign(1.0f / std::sqrt(fu * fu + f.b0 * f.b0), fu)));
    const float xu = std::min(f.x1, std::max(f.x0, -cu * f.z0 / std::sqrt(std::max(0.0f, 1.0f - cu * cu))));
    const float dd = std::sqrt(xu * xu + f.z0 * f.z0);
    const float h0 = f.y0 / std::sqrt(dd * dd + f.y0 * f.y0), h1 = f.y1 / std::sqrt(dd * dd + f.y1 * f.y1);
    const float hv = h0 + s1 * (h1 - h0);
    const float yv = std::min(f.y1, std::max(f.y0, hv * dd / std::sqrt(std::max(0.0f, 1.0f - hv * hv))));
    u = std::clamp((xu - f.x0) / (f.x1 - f.x0), 0.0f, 1.0f);
    v = std::clamp((yv - f.y0) / (f.y1 - f.y0), 0.0f, 1.0f);
}
// Rectangle coordinates -> square sample (the inverse of rect_sample).
static void rect_inverse(const RectFrame &f, float u, float v, float &s0, float &s1) {
    float xu = f.x0 + u * (f.x1 - f.x0);
    const float yv = f.y0 + v * (f.y1 - f.y0);
    if (xu == 0.0f) xu = 1e-10f;
    const float fusq = std::max(0.0f, 1.0f + f.z0 * f.z0 / (xu * xu) - f.b0 * f.b0);
    const float fu = std::copysign(std::sqrt(fusq), xu);
    const float root = std::sqrt(std::max(0.0f, f.b0 * f.b0 - f.b1 * f.b1 + fusq));
    float au = atan2_f(-f.b1 * fu - std::copysign(f.b0 * root, fu * f.b0), f.b0 * f.b1 - root * std::fabs(fu));
    if (au > 0) au = au - 2.0f * kPi;
    s0 = std::clamp((au + f.g2 + f.g3) / f.area, 0.0f, 1.0f);
    const float dd2 = xu * xu + f.z0 * f.z0;
    const float h0 = f.y0 / std::sqrt(dd2 + f.y0 * f.y0), h1 = f.y1 / std::sqrt(dd2 + f.y1 * f.y1);
    const float hv = yv / std::sqrt(dd2 + yv * yv);
    s1 = std::clamp((hv - h0) / (h1 - h0), 0.0f, 1.0f);
}
// Area density of a quad point q seen from r, as solid-angle density.
static float quad_area_pdf(const Light &L, V3 r, V3 q) {
    const V3 delta = q - r;
    const float d2 = dot(delta, delta);
    if (!(d2 > 0)) return 0.0f;
    const float c = std::fabs(dot(L.uz, delta * (1.0f / std::sqrt(d2))));
    return c > 0 ? L.inv_area * d2 / c : 0.0f;
}
// Solid-angle density of the quad proposal at rectangle coordinates (u, v)
// (point q on the quad) seen from r with receiver normal n.
static float quad_pdf(const Light &L, V3 r, V3 n, float u, float v, V3 q) {
    if (L.rect) {
        const RectFrame f = rect_frame(L, r);
        if (f.area > 1e-4f) {
            float s0, s1;
            rect_inverse(f, u, v, s0, s1);
            const bool has_n = (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f) && !slip(SLIP_RECT_PDF_UNWARPED);
            return (has_n ? bilinear_density(s0, s1, rect_cosines(f, L, n)) : 1.0f) / f.area;
        }
    }
    return quad_area_pdf(L, r, q);
}
static void quad_coords(const Light &L, V3 q, float &u, float &v) {
    const V3 d = q - L.p00;
    u = std::clamp(dot(d, L.ex) / dot(L.ex, L.ex), 0.0f, 1.0f);
    v = std::clamp(dot(d, L.ey) / dot(L.ey, L.ey), 0.0f, 1.0f);
}

// --- triangle emitters: the spherical-triangle proposal (Arvo 1995, as
// PBRT v4's SampleSphericalTriangle with the reference's rationalised
// 1 - cos(beta'); src/light_geometry_sampling_shared.h:109-176,
// shaders/light_geometry_sampling.h:114-209) for triangles subtending 3e-4
// to 6.22 sr, warped by the receiver's vertex cosines when a normal is
// given; otherwise uniform over the area, (b1, b2) = (1 - sqrt u0, u1 sqrt u0).
static bool positive(V3 v) { return dot(v, v) > 0; }
static float tri_solid_angle(V3 p0, V3 p1, V3 p2, V3 r) {
    const V3 a = p0 - r, b = p1 - r, c = p2 - r;
    if (!positive(a) || !positive(b) || !positive(c)) return 0.0f;
    return spherical_triangle_area(unit(a), unit(b), unit(c));
}
static bool spherical_tri(float area) { return area >= 3e-4f && area <= 6.22f; }
// Bilinear weights of the vertex cosines: (b, b, a, c) over the square.
static Bilinear tri_cosines(V3 p0, V3 p1, V3 p2, V3 r, V3 n) {
    const float wa = std::max(0.01f, std::fabs(dot(n, unit(p0 - r))));
    const float wb = std::max(0.01f, std::fabs(dot(n, unit(p1 - r))));
    const float wc = std::max(0.01f, std::fabs(dot(n, unit(p2 - r))));
    return {wb, wb, wa, wc};
}
// Square sample -> barycentrics (b1, b2) of p1 and p2 on the triangle seen
// from r.  False: a degenerate configuration.
static bool tri_sample(V3 p0, V3 p1, V3 p2, V3 r, float u0, float u1, float &b1, float &b2) {
    const V3 a = unit(p0 - r), b = unit(p1 - r), c = unit(p2 - r);
    const V3 nab = cross3(a, b), nca = cross3(c, a);
    if (!positive(nab) || !positive(nca)) return false;
    const float alpha = angle_between(unit(nab), -unit(nca));
    const float area = spherical_triangle_area(a, b, c);
    const float ca = cos_f(alpha), sa = sin_f(alpha);
    const float subarea = u0 * area;
    const float sh = sin_f(subarea / 2), st = sin_f(subarea), ct_area = cos_f(subarea);
    const float one_minus_ct = 2 * sh * sh;
    const float sp = sa * ct_area - ca * st;
    const float cp = -ca * ct_area - sa * st;
    const float k1 = -2 * sin_f(alpha - subarea / 2) * sh;
    const float k2 = sa * (0.5f * dot(a - b, a - b) - one_minus_ct) - ca * st;
    const float divisor = (k2 * sp + k1 * cp) * sa;
    if (divisor == 0) return false;
    const float omc = std::min(2.0f, std::max(0.0f, (-k2 * one_minus_ct - k1 * st) / divisor));
    const float cosbp = 1 - omc;
    const V3 arc = c - dot(c, a) * a;
    if (!positive(arc)) return false;
    const V3 cprime = cosbp * a + std::sqrt(std::max(0.0f, omc * (2 - omc))) * unit(arc);
    const float omct = u1 * 0.5f * dot(cprime - b, cprime - b);
    const float ct = 1 - omct;
    const V3 tangent = cprime - dot(cprime, b) * b;
    if (!positive(tangent)) return false;
    const V3 dir = ct * b + std::sqrt(std::max(0.0f, omct * (2 - omct))) * unit(tangent);
    const V3 e1 = p1 - p0, e2 = p2 - p0, s1 = cross3(dir, e2);
    const float denom = dot(s1, e1);
    if (denom == 0) return false;
    const V3 s = r - p0;
    b1 = std::min(1.0f, std::max(0.0f, dot(s, s1) / denom));
    b2 = std::min(1.0f, std::max(0.0f, dot(dir, cross3(s, e1)) / denom));
    if (b1 + b2 > 1) {
        const float sum = b1 + b2;
        b1 = b1 / sum;
        b2 = b2 / sum;
    }
    return std::isfinite(b1) && std::isfinite(b2);
}
// Direction dir from r (a point of the triangle) -> square sample.
static void tri_inverse(V3 p0, V3 p1, V3 p2, V3 r, V3 dir, float &s0, float &s1) {
    const V3 a = unit(p0 - r), b = unit(p1 - r), c = unit(p2 - r);
    V3 cp = unit(cross3(cross3(b, dir), cross3(c, a)));
    if (dot(cp, a + c) < 0) cp = -cp;
    s0 = std::min(1.0f, std::max(0.0f, spherical_triangle_area(a, b, cp) / spherical_triangle_area(a, b, c)));
    s1 = std::min(1.0f, std::max(0.0f, dot(dir - b, dir - b) / dot(cp - b, cp - b)));
}
// Solid-angle density of the triangle proposal at point q (geometric normal
// ng) from r with receiver normal n.
static float tri_pdf(const Light &L, V3 p0, V3 p1, V3 p2, V3 r, V3 n, V3 q, V3 ng) {
    const V3 dv = q - r;
    const float d2 = dot(dv, dv);
    if (!(d2 > 0)) return 0.0f;
    const float area = tri_solid_angle(p0, p1, p2, r);
    if (spherical_tri(area)) {
        float s0, s1;
        tri_inverse(p0, p1, p2, r, unit(dv), s0, s1);
        const bool has_n = (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f) && !slip(SLIP_RECT_PDF_UNWARPED);
        return (has_n ? bilinear_density(s0, s1, tri_cosines(p0, p1, p2, r, n)) : 1.0f) / area;
    }
    const float cl = std::fabs(dot(ng, dv / std::sqrt(d2)));
    return cl > 0 ? L.inv_area * d2 / cl : 0.0f;
}


// A light sample from point p (normal n; zero in media): selection by the
// scene's light sampler, then the light's own proposal.  pdf includes the
// selection probability; for delta lights it is that probability alone.
static LightSample sample_light(const Scene &s, V3 p, V3 n, const Wavelengths &wl, float ul, float u0, float u1) {
    LightSample ls;
    const Selected sel = light_select(s, p, n, ul);
    if (sel.light == RP_NONE || !(sel.pmf > 0)) return ls;
    const Light &L = s.lights[sel.light];
    if (L.kind == 0 || L.kind == 3) {
        float pdf_omega;
        uint32_t tri;
        if (L.kind == 0) {
            V3 a, b, c;
            s.tri(L.index, a, b, c);
            float su = std::min(u0, 0x1.fffffep-1f), sv = std::min(u1, 0x1.fffffep-1f), b1, b2;
            const float area = tri_solid_angle(a, b, c, p);
            bool spherical = false;
            if (spherical_tri(area)) {
                // the warp's bilinear density over the solid angle, at the warped sample
                float dens = 1.0f;
                if (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f) {
                    const Bilinear w = tri_cosines(a, b, c, p, n);
                    bilinear_warp(su, sv, w);
                    if (!slip(SLIP_RECT_PDF_UNWARPED)) dens = bilinear_density(su, sv, w);
                }
                if (!tri_sample(a, b, c, p, su, sv, b1, b2)) return ls;
                pdf_omega = dens / area;
                spherical = true;
            } else {
                const float root = std::sqrt(su);
                b1 = 1.0f - root;
                b2 = sv * root;
            }
            const float b0 = 1.0f - b1 - b2;
            ls.p = b0 * a + b1 * b + b2 * c;
            ls.err = gamma_n(6) * (vabs(b0 * a) + vabs(b1 * b) + vabs(b2 * c));
            ls.n = normalize(cross(a - c, b - c));
            tri = L.index;
            if (!spherical) {
                const V3 dv = ls.p - p;
                const float d2 = dot(dv, dv);
                const float cl = std::fabs(dot(ls.n, dv / std::sqrt(d2)));
                pdf_omega = d2 > 0 && cl > 0 ? L.inv_area * d2 / cl : 0.0f;
            }
        } else {
            float su = std::min(u0, 0x1.fffffep-1f), sv = std::min(u1, 0x1.fffffep-1f), qu = su, qv = sv;
            bool spherical = false;
            if (L.rect) {
                // spherical rectangle above 1e-4 sr: the density is the warp's
                // bilinear density over the solid angle, at the warped sample
                const RectFrame f = rect_frame(L, p);
                if (f.area > 1e-4f) {
                    float dens = 1.0f;
                    if (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f) {
                        const Bilinear w = rect_cosines(f, L, n);
                        bilinear_warp(su, sv, w);
                        if (!slip(SLIP_RECT_PDF_UNWARPED)) dens = bilinear_density(su, sv, w);
                    }
                    rect_sample(f, su, sv, qu, qv);
                    pdf_omega = dens / f.area;
                    spherical = true;
                }
            }
            const V3 a = qu * L.ex, b = qv * L.ey;
            ls.p = L.p00 + a + b;
            ls.err = gamma_n(6) * (vabs(L.p00) + vabs(a) + vabs(b));
            ls.n = L.uz;
            tri = L.index;
            if (!spherical) pdf_omega = quad_area_pdf(L, p, ls.p);
        }
        const V3 dv = ls.p - p;
        const float d2 = dot(dv, dv);
        if (!(d2 > 0)) return ls;
        ls.dist = std::sqrt(d2);
        ls.wi = dv / ls.dist;
        const float cos_l = dot(ls.n, -ls.wi);
        const auto &em = s.emissions[s.objects[s.tri_object[tri]].emission];
        if (!em.two_sided && !(cos_l > 0)) return ls;
        if (cos_l == 0.0f) return ls;
        ls.pdf = sel.pmf * pdf_omega;
        for (int k = 0; k < 4; ++k) ls.le[k] = spectrum_eval(em.radiance, wl.lambda[k], s.spectra.data());
        ls.ok = ls.pdf > 0 && std::isfinite(ls.pdf);
        return ls;
    }
    if (L.kind == 1) {
        const auto &sp = s.spots[L.index];
        V3 pos = {sp.position[0], sp.position[1], sp.position[2]};
        V3 dv = pos - p;
        float d2 = dot(dv, dv);
        if (!(d2 > 0)) return ls;
        ls.dist = std::sqrt(d2);
        ls.wi = dv / ls.dist;
        V3 axis = {sp.direction[0], sp.direction[1], sp.direction[2]};
        float cos_t = dot(-ls.wi, axis);
        float t = (cos_t - sp.cos_falloff_end) / (sp.cos_falloff_start - sp.cos_falloff_end);
        t = std::min(std::max(t, 0.0f), 1.0f);
        float fall = slip(SLIP_SPOT_NO_FALLOFF) ? 1.0f : t * t * (3.0f - 2.0f * t);
        for (int k = 0; k < 4; ++k)
            ls.le[k] = spectrum_eval(sp.intensity, wl.lambda[k], s.spectra.data()) * fall / d2;
        ls.p = pos;
        ls.n = {0, 0, 0};
        ls.err = {0, 0, 0};
        ls.pdf = sel.pmf;
        ls.delta = true;
        ls.ok = true;
        return ls;
    }
    // environment: uniform sphere
    float z = 1.0f - 2.0f * u0;
    float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    float sn, cs;
    sincos_turns(u1, sn, cs);
    ls.wi = {r * cs, r * sn, z};
    ls.infinite = true;
    ls.pdf = sel.pmf * kInv4Pi;
    for (int k = 0; k < 4; ++k) ls.le[k] = spectrum_eval(s.d.environment, wl.lambda[k], s.spectra.data());
    ls.ok = true;
    return ls;
}

// Solid-angle density of light sampling from (p, n) toward the emitter point
// q of light `id` (an emission hit reached by BSDF sampling): the selection
// probability times the light's proposal density.
static float area_light_pdf(const Scene &s, uint32_t id, V3 p, V3 n, V3 q, V3 ng) {
    const float pmf = slip(SLIP_LIGHT_PMF_DROPPED)   ? 1.0f
                      : slip(SLIP_BVH_PMF_NO_NORMAL) ? light_pmf(s, p, V3{0, 0, 0}, id)
                                                     : light_pmf(s, p, n, id);
    const Light &L = s.lights[id];
    if (L.kind == 3) {
        float u, v;
        quad_coords(L, q, u, v);
        return pmf * quad_pdf(L, p, n, u, v, q);
    }
    V3 a, b, c;
    s.tri(L.index, a, b, c);
    return pmf * tri_pdf(L, a, b, c, p, n, q, ng);
}

static float env_pdf(const Scene &s) {
    if (s.env_light == RP_NONE) return 0.0f;
    return s.env_pmf * (slip(SLIP_ENV_PDF_HEMISPHERE) ? 2.0f * kInv4Pi : kInv4Pi);
}

// Forward-mode derivative of the detached estimator (independent of the
// device's reverse-mode replay): d log(beta_k) / d theta_p is carried along
// the path and every contribution C_k adds a_k * C_k * d log C_k / d theta_p.
struct NoTrack {
    static constexpr bool on = false;
    void begin(const Wavelengths &) {}
    void medium_factor(uint32_t, const Wavelengths &, float, const float *) {}
    void medium_pass(uint32_t, const Wavelengths &, float) {}
    void add(const float *) {}
    void save_direct() {}
    void continuation(const rp_material &, uint32_t, const Lane *, V3, V3, const Wavelengths &) {}
    void nee(const rp_material &, uint32_t, const Lane *, V3, V3, const Wavelengths &, const float *, float) {}
    void nee_medium(const float *, float) {}
};

struct GradSpec {
    std::vector<int32_t> refl_base, rough, sigma_base;
    std::vector<uint32_t> refl_bins, sigma_bins;
    uint32_t nparams = 0;
    uint32_t sigma_medium = RP_NONE;
};

static uint32_t bin_of(float lambda, uint32_t bins) {
    int i = int(std::floor(lambda + 0.5f)) - 360;
    i = std::min(std::max(i, 0), 470);
    return std::min(bins - 1, uint32_t(i) * bins / 471u);
}

// d log f / d r of an isotropic rough conductor, alpha = sqrt(r) (remapped) or r.
static double conductor_dlogf(const Lane &l, const rp_material &m, V3 wo, V3 wi) {
    if (l.smooth || !(wo.z * wi.z > 0)) return 0.0;
    double r = m.roughness_u;
    double a_raw = m.remap_roughness ? std::sqrt(r) : r;
    if (a_raw < 1e-4) return 0.0;
    double a = l.ax;
    double hx = double(wo.x) + wi.x, hy = double(wo.y) + wi.y, hz = double(wo.z) + wi.z;
    double hn = std::sqrt(hx * hx + hy * hy + hz * hz);
    hx /= hn, hy /= hn, hz /= hn;
    // D = 1 / (pi a^2 (s/a^2 + c^2)^2),  s = sin^2, c = cos of the half vector
    double sn = hx * hx + hy * hy, t = sn / (a * a) + hz * hz;
    double dlogD = -2.0 / a + 4.0 * sn / (a * a * a * t);
    auto lam = [&](V3 w, double &dl) {
        double tan2 = (double(w.x) * w.x + double(w.y) * w.y) / (double(w.z) * w.z);
        double root = std::sqrt(1.0 + a * a * tan2);
        dl = 0.5 * a * tan2 / root;
        return 0.5 * (root - 1.0);
    };
    double dlo, dli;
    double lo = lam(wo, dlo), li = lam(wi, dli);
    double dlogf = dlogD - (dlo + dli) / (1.0 + lo + li);
    return dlogf * (m.remap_roughness ? 0.5 / a : 1.0);
}

struct ForwardGrad {
    static constexpr bool on = true;
    const GradSpec *g = nullptr;
    double a[4] = {0, 0, 0, 0};
    std::vector<double> dlb, ddlb; // [4][P] current and pre-continuation log-derivatives
    std::vector<double> *grad = nullptr;
    // pending NEE local derivative
    int32_t nee_idx[4];
    double nee_dlog[4];
    uint32_t P() const { return g->nparams; }
    void begin(const Wavelengths &) {
        dlb.assign(4 * size_t(P()), 0.0);
        ddlb = dlb;
    }
    void bump(int k, int32_t idx, double v) {
        if (idx >= 0) dlb[size_t(k) * P() + idx] += v;
    }
    void medium_factor(uint32_t m, const Wavelengths &wl, float t, const float *ss) {
        if (g->sigma_base[m] < 0) return;
        for (int k = 0; k < 4; ++k)
            if (ss[k] > 0) bump(k, g->sigma_base[m] + int32_t(bin_of(wl.lambda[k], g->sigma_bins[m])), 1.0 / ss[k] - t);
    }
    void medium_pass(uint32_t m, const Wavelengths &wl, float d) {
        if (g->sigma_base[m] < 0) return;
        for (int k = 0; k < 4; ++k) bump(k, g->sigma_base[m] + int32_t(bin_of(wl.lambda[k], g->sigma_bins[m])), -double(d));
    }
    void add(const float *v) { // contribution carried by the current throughput
        for (int k = 0; k < 4; ++k)
            if (v[k] != 0.0f)
                for (uint32_t p = 0; p < P(); ++p) (*grad)[p] += a[k] * double(v[k]) * dlb[size_t(k) * P() + p];
    }
    void save_direct() { ddlb = dlb; }
    void local(const rp_material &mat, uint32_t mid, const Lane *lanes, V3 wo, V3 wi, const Wavelengths &wl,
               int32_t idx[4], double dl[4]) {
        for (int k = 0; k < 4; ++k) idx[k] = -1, dl[k] = 0.0;
        if (mat.kind == RP_MATERIAL_DIFFUSE && g->refl_base[mid] >= 0) {
            for (int k = 0; k < 4; ++k)
                if (lanes[k].r > 0 && wo.z * wi.z > 0) {
                    idx[k] = g->refl_base[mid] + int32_t(bin_of(wl.lambda[k], g->refl_bins[mid]));
                    dl[k] = 1.0 / lanes[k].r;
                }
        } else if (mat.kind == RP_MATERIAL_CONDUCTOR && g->rough[mid] >= 0) {
            for (int k = 0; k < 4; ++k) {
                idx[k] = g->rough[mid];
                dl[k] = conductor_dlogf(lanes[k], mat, wo, wi);
            }
        }
    }
    void continuation(const rp_material &mat, uint32_t mid, const Lane *lanes, V3 wo, V3 wi, const Wavelengths &wl) {
        int32_t idx[4];
        double dl[4];
        local(mat, mid, lanes, wo, wi, wl, idx, dl);
        for (int k = 0; k < 4; ++k) bump(k, idx[k], dl[k]);
    }
    // NEE contribution from a surface: pre-continuation derivatives + BSDF at the light direction + shadow medium.
    void nee(const rp_material &mat, uint32_t mid, const Lane *lanes, V3 wo, V3 wi, const Wavelengths &wl,
             const float *c, float mlen) {
        int32_t idx[4];
        double dl[4];
        local(mat, mid, lanes, wo, wi, wl, idx, dl);
        finish_nee(wl, c, mlen, idx, dl, ddlb);
    }
    void nee_medium(const float *c, float mlen) {
        int32_t idx[4] = {-1, -1, -1, -1};
        double dl[4] = {0, 0, 0, 0};
        Wavelengths wl = cur_wl;
        finish_nee(wl, c, mlen, idx, dl, dlb);
    }
    void finish_nee(const Wavelengths &wl, const float *c, float mlen, const int32_t idx[4], const double dl[4],
                    const std::vector<double> &base) {
        for (int k = 0; k < 4; ++k) {
            if (c[k] == 0.0f) continue;
            double ak = a[k] * double(c[k]);
            for (uint32_t p = 0; p < P(); ++p) (*grad)[p] += ak * base[size_t(k) * P() + p];
            if (idx[k] >= 0) (*grad)[idx[k]] += ak * dl[k];
            if (g->sigma_medium != RP_NONE && mlen > 0) {
                int32_t si = g->sigma_base[g->sigma_medium] + int32_t(bin_of(wl.lambda[k], g->sigma_bins[g->sigma_medium]));
                (*grad)[si] += ak * (-double(mlen));
            }
        }
    }
    Wavelengths cur_wl{};
};

// One complete hero path: src/render.metal:238-312 (begin_sample, primary_ray)
// followed by the event loop of shaders/packet_pt_event_body.inc.
template <class T>
static PathOut trace_path_t(const Ctx &c, uint32_t pixel, uint64_t sample, T &trk);
static PathOut trace_path(const Ctx &c, uint32_t pixel, uint64_t sample) {
    NoTrack nt;
    return trace_path_t(c, pixel, sample, nt);
}
// Test switch (or_set_roulette): with roulette off, a path's sampling no
// longer depends on its throughput, so finite differences of the renderer
// with common random numbers are smooth in reflectance parameters and pin
// the forward-mode derivative exactly.  The GPU renderer has no such switch.
static std::atomic<bool> g_roulette{true};

template <class T>
static PathOut trace_path_t(const Ctx &c, uint32_t pixel, uint64_t sample, T &trk) {
    const Scene &s = *c.s;
    const Scene &s0 = s.sampling ? *s.sampling : s;
    PathOut out;
    const uint32_t W = s.d.camera.width, H = s.d.camera.height;
    const uint32_t px = pixel % W, py = pixel / W;
    uint32_t seed_bits = uint32_t(c.seed);
    out.index = sobol_index(s.m, px, py, sample);
    uint32_t w0 = sobol_word(out.index, 0, RP_RANDOMIZE_NONE, 0);
    uint32_t w1 = sobol_word(out.index, 1, RP_RANDOMIZE_NONE, 0);
    uint32_t w2 = sobol_word(out.index, 2, c.mode, seed_bits);
    out.w0 = w0;
    out.w1 = w1;
    Wavelengths wl = sample_wavelengths(sobol_uniform(w2), s.d.hero_wavelengths);
    out.wl = wl;
    trk.begin(wl);
    if constexpr (T::on) trk.cur_wl = wl;
    uint32_t mlen_medium = RP_NONE;
    if constexpr (T::on) mlen_medium = trk.g->sigma_medium;
    float ux = sobol_camera_uniform(w0, s.m, px), uy = sobol_camera_uniform(w1, s.m, py);
    float fx = (float(px) + ux) / float(W), fy = (float(py) + uy) / float(H);
    const rp_camera &cam = s.d.camera;
    V3 d = {cam.corner[0] + fx * cam.dx[0] + fy * cam.dy[0], cam.corner[1] + fx * cam.dx[1] + fy * cam.dy[1],
            cam.corner[2] + fx * cam.dx[2] + fy * cam.dy[2]};
    float scale = std::max(std::fabs(d.x), std::max(std::fabs(d.y), std::fabs(d.z)));
    d = normalize(d / scale);
    V3 o = {cam.origin[0], cam.origin[1], cam.origin[2]};
    float tmax = INFINITY;

    float beta[4] = {1, 1, 1, 1}, density[4] = {1, 1, 1, 1};
    float prev_density[4] = {1, 1, 1, 1}, prev_pdf[4] = {0, 0, 0, 0};
    float eta_scale[4] = {1, 1, 1, 1};
    bool rr_enabled = g_roulette.load(std::memory_order_relaxed), prev_delta = false;
    V3 prev_p = o, prev_n = {0, 0, 0};
    uint32_t count = 1, crossings = 0, medium = cam.medium, mdraw = 0;
    const uint32_t capacity = s.d.max_depth + 2;
    auto add = [&](const float v[4]) {
        for (int k = 0; k < 4; ++k) out.L[k] = out.L[k] + v[k];
    };

    for (;;) {
        if (count >= capacity) break;
        if (!(beta[0] > 0 || beta[1] > 0 || beta[2] > 0 || beta[3] > 0)) break;
        Hit h = closest(s, o, d, tmax);
        ++out.ext_rays;
        if (count == 1 && crossings == 0) out.primary = h.tri;
        float seg = h.tri == RP_NONE ? INFINITY : h.t;

        // ---- homogeneous medium along the segment (hero free flight)
        if (medium != RP_NONE) {
            MediumCoeff mc = medium_coeff(s, medium, wl);
            const MediumCoeff mc0 = &s0 == &s ? mc : medium_coeff(s0, medium, wl);
            auto mw = medium_draws(c, pixel, sample, mdraw);
            float uf = philox_u01(mw[0]);
            float tf = mc0.st[0] > 0 ? -log_f(1.0f - uf) / mc0.st[0] : INFINITY;
            if (tf < seg) {
                // scattering event in the medium at distance tf
                float num[4], pdf[4];
                for (int k = 0; k < 4; ++k) {
                    float tr = exp_f(-mc.st[k] * tf);
                    num[k] = mc.ss[k] * tr;
                    pdf[k] = mc0.st[k] * exp_f(-mc0.st[k] * tf);
                }
                if (!advance(density, beta, num, pdf)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
                trk.medium_factor(medium, wl, tf, mc.ss);
                for (int k = 0; k < 4; ++k) prev_pdf[k] = prev_pdf[k] * pdf[k];
                V3 mp = o + tf * d;
                ++count;
                crossings = 0;
                if (count == capacity) break;
                if (count - 2 >= c.emergency) { out.status = RP_PATH_EMERGENCY_DEPTH; break; }
                auto dr = draws(c, out.index, 4 + count - 2);
                float direct_density[4], direct_beta[4];
                std::memcpy(direct_density, density, sizeof density);
                bool direct_allowed = hero_continue(beta, eta_scale, rr_enabled, count - 1, sobol_uniform(dr[3]));
                std::memcpy(direct_beta, beta, sizeof beta);
                V3 wo = -d;
                if (!direct_allowed) break;
                // phase sampling (hero g is wavelength independent)
                float ppdf;
                V3 wi = hg_sample(wo, mc.g, sobol_uniform(dr[1]), sobol_uniform(dr[2]), ppdf);
                std::memcpy(prev_density, density, sizeof density);
                float pnum[4] = {ppdf, ppdf, ppdf, ppdf}, ppdfs[4] = {ppdf, ppdf, ppdf, ppdf};
                if (!advance(density, beta, pnum, ppdfs)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
                for (int k = 0; k < 4; ++k) prev_pdf[k] = ppdf;
                prev_delta = false;
                prev_p = mp;
                prev_n = {0, 0, 0};
                // NEE from the medium vertex
                auto lw = draws(c, out.index, 133 + count - 2);
                LightSample ls = sample_light(s, mp, V3{0, 0, 0}, wl, sobol_uniform(lw[0]), sobol_uniform(lw[1]),
                                              sobol_uniform(lw[2]));
                if (ls.ok) {
                    float ph = hg(slip(SLIP_PHASE_SIGN) ? -dot(wo, ls.wi) : dot(wo, ls.wi), mc.g);
                    float val[4];
                    bool any = false;
                    for (int k = 0; k < 4; ++k) {
                        val[k] = ((direct_beta[k] * ph) * (ls.le[k] / ls.pdf));
                        any = any || val[k] > 0;
                    }
                    if (any) {
                        float tr[4];
                        uint32_t cr = 0;
                        V3 to;
                        float tm;
                        V3 sd;
                        if (ls.infinite) { sd = ls.wi; tm = INFINITY; }
                        else {
                            to = ls.delta ? ls.p : offset_origin(ls.p, ls.err, ls.n, mp - ls.p);
                            sd = to - mp;
                            tm = 1.0f - 1e-4f;
                        }
                        ++out.shadow_rays;
                        float mlen = 0.0f;
                        if (transmittance(c, mp, sd, tm, medium, wl, tr, cr, &mlen, mlen_medium)) {
                            if (!slip(SLIP_NEE_NO_TRANSMITTANCE))
                                for (int k = 0; k < 4; ++k) val[k] = val[k] * tr[k];
                            if (!ls.delta) {
                                float lp[4] = {ls.pdf, ls.pdf, ls.pdf, ls.pdf};
                                float pp[4] = {ph, ph, ph, ph};
                                float a = mixture(direct_density, lp), b = mixture(direct_density, pp);
                                float w = a / (a + b);
                                for (int k = 0; k < 4; ++k) val[k] = val[k] * w;
                            }
                            add(val);
                            trk.nee_medium(val, mlen);
                        }
                    }
                }
                o = mp;
                d = wi;
                tmax = INFINITY;
                continue;
            }
            // passes to the surface (or escapes): transmission probability per lane
            if (std::isfinite(seg)) {
                float num[4], pdf[4];
                for (int k = 0; k < 4; ++k) {
                    float tr = exp_f(-mc.st[k] * seg);
                    num[k] = slip(SLIP_HERO_TRANSMITTANCE) ? exp_f(-mc.st[0] * seg) : tr;
                    pdf[k] = &s0 == &s ? tr : exp_f(-mc0.st[k] * seg);
                }
                if (!advance(density, beta, num, pdf)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
                trk.medium_pass(medium, wl, seg);
                for (int k = 0; k < 4; ++k) prev_pdf[k] = prev_pdf[k] * pdf[k];
            } else {
                break; // infinite medium, nothing reached
            }
        }

        if (h.tri == RP_NONE) {
            if (s.d.has_environment) {
                float val[4];
                for (int k = 0; k < 4; ++k)
                    val[k] = beta[k] * spectrum_eval(s.d.environment, wl.lambda[k], s.spectra.data());
                if (!(count == 1 || prev_delta)) {
                    float lp = env_pdf(s);
                    float lpv[4] = {lp, lp, lp, lp};
                    float a = mixture(prev_density, prev_pdf), b = mixture(prev_density, lpv);
                    float w = a / (a + b);
                    for (int k = 0; k < 4; ++k) val[k] = val[k] * w;
                }
                add(val);
                trk.add(val);
            }
            break;
        }
        SurfacePoint sp = surface_point(s, h);
        const rp_object &obj = s.objects[s.tri_object[h.tri]];
        const rp_material &mat = s.materials[obj.material];
        if (mat.kind == RP_MATERIAL_INTERFACE) {
            if (crossings++ >= c.boundary) { out.status = RP_PATH_BOUNDARY_EXHAUSTION; break; }
            medium = outgoing_medium(obj, medium, sp.n, d);
            o = offset_origin(sp.p, sp.err, sp.n, d);
            tmax = INFINITY;
            continue;
        }
        V3 wo = -d;
        ++count;
        crossings = 0;
        // arrival emission (hero_endpoint)
        if (obj.emission != RP_NONE && s.tri_light[h.tri] != RP_NONE) {
            const auto &em = s.emissions[obj.emission];
            float cos_l = dot(sp.n, wo);
            if (em.two_sided || cos_l > 0) {
                float val[4];
                for (int k = 0; k < 4; ++k) val[k] = beta[k] * spectrum_eval(em.radiance, wl.lambda[k], s.spectra.data());
                if (!(count == 2 || prev_delta) && !slip(SLIP_EMISSION_UNWEIGHTED)) {
                    float lp = area_light_pdf(s, s.tri_light[h.tri], prev_p, prev_n, sp.p, sp.n);
                    float lpv[4] = {lp, lp, lp, lp};
                    float a = mixture(prev_density, prev_pdf), b = mixture(prev_density, lpv);
                    float w = a / (a + b);
                    for (int k = 0; k < 4; ++k) val[k] = val[k] * w;
                }
                add(val);
                trk.add(val);
            }
        }
        if (count == capacity) break;
        if (count - 2 >= c.emergency) { out.status = RP_PATH_EMERGENCY_DEPTH; break; }
        Frame fr = frame_from_z(sp.n);
        V3 wol = fr.local(wo);
        Lane lanes[4], lanes0[4];
        for (int k = 0; k < 4; ++k) lanes[k] = prepare_lane(s, mat, wl.lambda[k]);
        for (int k = 0; k < 4; ++k)
            lanes0[k] = &s0 == &s ? lanes[k] : prepare_lane(s0, s0.materials[obj.material], wl.lambda[k]);
        bool connectible = !lanes0[0].smooth || lanes0[0].kind == RP_MATERIAL_DIFFUSE;
        auto dr = draws(c, out.index, 4 + count - 2);
        float direct_density[4], direct_beta[4];
        std::memcpy(direct_density, density, sizeof density);
        bool direct_allowed = hero_continue(beta, eta_scale, rr_enabled, count - 1, sobol_uniform(dr[3]));
        std::memcpy(direct_beta, beta, sizeof beta);
        trk.save_direct();
        bool cont = direct_allowed;
        if (cont) {
            float uc = sobol_uniform(dr[0]), u0 = sobol_uniform(dr[1]), u1 = sobol_uniform(dr[2]);
            Sampled hs = sample_lane(lanes0[0], wol, uc, u0, u1);
            if (!hs.ok) cont = false;
            else {
                float num[4], pdf[4], etas[4];
                if (hs.delta) {
                    for (int k = 0; k < 4; ++k) {
                        float m, p;
                        dirac_lane(lanes[k], wol, hs.wi, uc, u0, u1, m, p);
                        num[k] = m;
                        pdf[k] = p;
                        if (&s0 != &s) dirac_lane(lanes0[k], wol, hs.wi, uc, u0, u1, m, pdf[k]);
                        if (k == 0) pdf[0] = hs.pdf;
                        etas[k] = hs.transmission ? (wol.z > 0 ? lanes[k].eta : 1.0f / lanes[k].eta) : 1.0f;
                    }
                } else {
                    for (int k = 0; k < 4; ++k) {
                        Eval e = eval_lane(lanes[k], wol, hs.wi);
                        num[k] = slip(SLIP_CONTINUATION_NO_COSINE) ? e.f : e.f * std::fabs(hs.wi.z);
                        pdf[k] = &s0 == &s ? e.pdf : eval_lane(lanes0[k], wol, hs.wi).pdf;
                        etas[k] = hs.transmission ? (wol.z > 0 ? lanes[k].eta : 1.0f / lanes[k].eta) : 1.0f;
                    }
                    pdf[0] = hs.pdf;
                }
                std::memcpy(prev_density, direct_density, sizeof density);
                if (!advance(density, beta, num, pdf)) {
                    out.status = RP_PATH_INVALID_ARITHMETIC;
                    cont = false;
                } else {
                    if (!hs.delta) trk.continuation(mat, obj.material, lanes, wol, hs.wi, wl);
                    std::memcpy(prev_pdf, pdf, sizeof pdf);
                    if (hs.transmission && rr_enabled)
                        for (int k = 0; k < 4; ++k) eta_scale[k] = eta_scale[k] * (etas[k] * etas[k]);
                    prev_delta = hs.delta;
                    prev_p = sp.p;
                    prev_n = oct_decode(oct_encode(sp.n));
                    V3 wi = fr.world(hs.wi);
                    wi = normalize(wi);
                    medium = outgoing_medium(obj, medium, sp.n, wi);
                    o = offset_origin(sp.p, sp.err, sp.n, wi);
                    d = wi;
                    tmax = INFINITY;
                }
            }
        }
        // direct lighting (hero_direct) with the pre-sampling state
        if (out.status == RP_PATH_SUCCESS && direct_allowed && connectible) {
            auto lw = draws(c, out.index, 133 + count - 2);
            LightSample ls = sample_light(s, sp.p, oct_decode(oct_encode(sp.n)), wl, sobol_uniform(lw[0]),
                                          sobol_uniform(lw[1]), sobol_uniform(lw[2]));
            if (ls.ok) {
                V3 wil = fr.local(ls.wi);
                float cosv = slip(SLIP_NEE_NO_COSINE) ? 1.0f : std::fabs(dot(sp.n, ls.wi));
                float val[4], bp[4];
                bool any = false;
                for (int k = 0; k < 4; ++k) {
                    Eval e = eval_lane(lanes[k], wol, wil);
                    bp[k] = &s0 == &s ? e.pdf : eval_lane(lanes0[k], wol, wil).pdf;
                    val[k] = ((direct_beta[k] * e.f) * (ls.le[k] / ls.pdf)) * cosv;
                    any = any || val[k] > 0;
                }
                if (any) {
                    V3 from = offset_origin(sp.p, sp.err, sp.n, ls.wi);
                    V3 sd;
                    float tm;
                    if (ls.infinite) { sd = ls.wi; tm = INFINITY; }
                    else {
                        V3 to = ls.delta ? ls.p : offset_origin(ls.p, ls.err, ls.n, from - ls.p);
                        sd = to - from;
                        tm = 1.0f - 1e-4f;
                    }
                    float tr[4];
                    uint32_t cr = 0;
                    uint32_t nm = outgoing_medium(obj, medium, sp.n, ls.wi);
                    ++out.shadow_rays;
                    float mlen = 0.0f;
                    if (transmittance(c, from, sd, tm, nm, wl, tr, cr, &mlen, mlen_medium)) {
                        for (int k = 0; k < 4; ++k) val[k] = val[k] * tr[k];
                        if (!ls.delta) {
                            float lp[4] = {ls.pdf, ls.pdf, ls.pdf, ls.pdf};
                            float a = mixture(direct_density, lp), b = mixture(direct_density, bp);
                            float w = a / (a + b);
                            for (int k = 0; k < 4; ++k) val[k] = val[k] * w;
                        }
                        add(val);
                        trk.nee(mat, obj.material, lanes, wol, wil, wl, val, mlen);
                    }
                }
            }
        }
        if (!cont) break;
    }
    out.vertices = count;
    return out;
}

// ============================================================ scalar transport
// One wavelength per path: the reference's scalar estimator for diffuse
// surfaces (shaders/scalar_transport.h:446-899; docs/rendering.md:41-60),
// restated.  Draw blocks (render.metal:154-177 draw_words): block 0 camera
// x/y and wavelength, scattering b uses block 1 + 2b (continuation x/y,
// roulette word z) and block 2 + 2b (light selection and position).  The
// independent sampler is Philox4x32-10 with counter (pixel, sample low,
// sample high, block) and key (seed low, seed high); Sobol blocks are the
// four dimensions 4 block .. 4 block + 3.
struct ScalarOut {
    uint32_t status = RP_PATH_SUCCESS;
    float wavelength = 0, radiance = 0;
    uint32_t scatterings = 0;
    uint32_t words[4] = {0, 0, 0, 0};
    uint32_t primary = RP_NONE;
    uint64_t ext_rays = 0, shadow_rays = 0;
};
// Power heuristic a^2 / (a^2 + b^2) on values scaled by their maximum
// (render.metal:1901-1909).
static float power_heuristic(float a, float b) {
    const float scale = std::max(a, b);
    if (!(scale > 0) || !std::isfinite(scale)) return NAN;
    a = a / scale;
    b = b / scale;
    return a * a / (a * a + b * b);
}
// Uniform sphere and cosine hemisphere about n (render.metal:821-836).
static V3 sphere_dir(float u, float v) {
    const float z = 1.0f - 2.0f * u;
    const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    float sn, cs;
    sincos_turns(v, sn, cs);
    return {r * cs, r * sn, z};
}
static V3 hemisphere_dir(V3 n, float u, float v) {
    const float r = std::sqrt(u);
    float sn, cs;
    sincos_turns(v, sn, cs);
    const V3 axis = std::fabs(n.z) < 0.999f ? V3{0, 0, 1} : V3{0, 1, 0};
    const V3 t = normalize(cross3(axis, n));
    const V3 b = cross3(n, t);
    return normalize((r * cs) * t + (r * sn) * b + std::sqrt(1.0f - u) * n);
}
// Demo MIS weights with p_light = 1/(4 pi), p_bsdf = cos/pi (render.metal:838-848).
static float demo_light_weight(float c) {
    const float q = 4.0f * c;
    return 1.0f / (1.0f + q * q);
}
static float demo_bsdf_weight(float c) {
    const float q = 4.0f * c;
    const float q2 = q * q;
    return q2 / (1.0f + q2);
}
// Light-sampling factor f cos / p_light times the power-heuristic weight
// (none for delta lights) (render.metal:2212-2219).
static float scattering_direct(float f_cos, float bsdf_pdf, float light_pdf, bool delta) {
    if (!(f_cos >= 0) || !(bsdf_pdf >= 0) || !(light_pdf > 0) || !std::isfinite(f_cos) || !std::isfinite(bsdf_pdf) ||
        !std::isfinite(light_pdf))
        return NAN;
    if (f_cos == 0) return 0;
    return (f_cos / light_pdf) * (delta ? 1.0f : power_heuristic(light_pdf, bsdf_pdf));
}
// Demo film weight: CIE linearly interpolated at lambda times 470 / Y
// (src/render.cpp:2131-2145; Y = the trapezoid integral of y-bar).
static void demo_sensor(float lambda, float L, float xyz[3]) {
    const auto &cie = cie_table();
    const float off = lambda - 360.0f;
    const int lo = std::min(std::max(int(off), 0), 470);
    const int hi = std::min(lo + 1, 470);
    const float f = off - float(lo);
    const float scale = 470.0f / 106.856914916766900f;
    for (int c = 0; c < 3; ++c) xyz[c] = ((cie.xyz[lo][c] + (cie.xyz[hi][c] - cie.xyz[lo][c]) * f) * scale) * L;
}

static ScalarOut trace_scalar(const Ctx &c, uint32_t pixel, uint64_t sample, bool independent) {
    const Scene &s = *c.s;
    ScalarOut out;
    const bool demo = s.d.transport == RP_TRANSPORT_SCALAR_DEMO;
    const uint32_t W = s.d.camera.width, H = s.d.camera.height;
    const uint32_t px = pixel % W, py = pixel / W;
    const uint32_t seed_lo = uint32_t(c.seed), seed_hi = uint32_t(c.seed >> 32);
    uint64_t index = 0;
    auto block = [&](uint32_t b) -> std::array<uint32_t, 4> {
        if (independent) return philox({pixel, uint32_t(sample), uint32_t(sample >> 32), b}, {seed_lo, seed_hi});
        return draws(c, index, b);
    };
    auto su = [&](uint32_t w) { return independent ? philox_u01(w) : sobol_uniform(w); };
    std::array<uint32_t, 4> w0;
    if (independent) w0 = block(0);
    else {
        index = sobol_index(s.m, px, py, sample);
        w0 = {sobol_word(index, 0, RP_RANDOMIZE_NONE, 0), sobol_word(index, 1, RP_RANDOMIZE_NONE, 0),
              sobol_word(index, 2, c.mode, seed_lo), sobol_word(index, 3, c.mode, seed_lo)};
    }
    for (int k = 0; k < 4; ++k) out.words[k] = w0[k];
    const float lambda = 360.0f + 470.0f * su(w0[2]);
    out.wavelength = lambda;
    Wavelengths wl;
    for (int k = 0; k < 4; ++k) wl.lambda[k] = lambda, wl.pdf[k] = 1.0f / 470.0f;
    const float ux = independent ? su(w0[0]) : sobol_camera_uniform(w0[0], s.m, px);
    const float uy = independent ? su(w0[1]) : sobol_camera_uniform(w0[1], s.m, py);
    const float fx = (float(px) + ux) / float(W), fy = (float(py) + uy) / float(H);
    const rp_camera &cam = s.d.camera;
    V3 d = {cam.corner[0] + fx * cam.dx[0] + fy * cam.dy[0], cam.corner[1] + fx * cam.dx[1] + fy * cam.dy[1],
            cam.corner[2] + fx * cam.dx[2] + fy * cam.dy[2]};
    const float scale = std::max(std::fabs(d.x), std::max(std::fabs(d.y), std::fabs(d.z)));
    d = normalize(d / scale);
    V3 o = {cam.origin[0], cam.origin[1], cam.origin[2]};
    const float E = s.d.has_environment ? spectrum_eval(s.d.environment, lambda, s.spectra.data()) : 0.0f;
    float throughput = 1.0f, escape_weight = 1.0f, prev_pdf = 0.0f;
    V3 prev_p = o, prev_n = {0, 0, 0};
    uint32_t crossings = 0;
    float L = 0.0f;
    for (;;) {
        Hit h = closest(s, o, d, INFINITY);
        ++out.ext_rays;
        if (out.scatterings == 0 && crossings == 0 && out.primary == RP_NONE) out.primary = h.tri;
        if (h.tri == RP_NONE) {
            if (!demo && out.scatterings != 0)
                escape_weight = power_heuristic(prev_pdf, s.env_light == RP_NONE ? 0.0f
                                                          : light_pmf(s, prev_p, prev_n, s.env_light) * kInv4Pi);
            L = L + throughput * escape_weight * E;
            break;
        }
        const uint32_t bounce = out.scatterings;
        const SurfacePoint sp = surface_point(s, h);
        const rp_object &obj = s.objects[s.tri_object[h.tri]];
        const rp_material &mat = s.materials[obj.material];
        if (mat.kind == RP_MATERIAL_INTERFACE) {
            if (crossings++ >= c.boundary) { out.status = RP_PATH_BOUNDARY_EXHAUSTION; break; }
            o = offset_origin(sp.p, sp.err, sp.n, d);
            continue;
        }
        if (obj.emission != RP_NONE && s.tri_light[h.tri] != RP_NONE) {
            const auto &em = s.emissions[obj.emission];
            const float Le = spectrum_eval(em.radiance, lambda, s.spectra.data());
            if (Le > 0.0f && (em.two_sided || dot(sp.n, -d) >= 0.0f)) {
                float weight = 1.0f;
                if (bounce != 0)
                    weight = power_heuristic(prev_pdf, area_light_pdf(s, s.tri_light[h.tri], prev_p, prev_n, sp.p, sp.n));
                if (!std::isfinite(weight)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
                L = L + throughput * (weight * Le);
            }
        }
        if (!demo && bounce >= s.d.max_depth) break;
        if (mat.kind != RP_MATERIAL_DIFFUSE) { out.status = RP_PATH_INVALID_HIT; break; }
        const float rho = spectrum_eval(mat.reflectance, lambda, s.spectra.data());
        if (!std::isfinite(rho) || !(rho >= 0.0f && rho <= 1.0f)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
        if (rho == 0.0f) break;
        if (bounce >= c.emergency) { out.status = RP_PATH_EMERGENCY_DEPTH; break; }
        const V3 n = dot(sp.n, d) > 0.0f ? -sp.n : sp.n;
        const auto lw = block(2 + 2 * bounce), bw = block(1 + 2 * bounce);
        if (demo) {
            const V3 ldir = sphere_dir(su(lw[0]), su(lw[1]));
            const float lc = dot(n, ldir);
            if (E > 0.0f && lc > 0.0f) {
                float tr[4];
                uint32_t cr = 0;
                ++out.shadow_rays;
                if (transmittance(c, offset_origin(sp.p, sp.err, sp.n, ldir), ldir, INFINITY, RP_NONE, wl, tr, cr))
                    L = L + throughput * rho * ((4.0f * lc) * demo_light_weight(lc)) * E * tr[0];
            }
        } else {
            const LightSample ls = sample_light(s, sp.p, sp.n, wl, su(lw[0]), su(lw[1]), su(lw[2]));
            if (ls.ok && ls.pdf > 0 && ls.le[0] > 0) {
                const float cosv = dot(n, ls.wi);
                if (cosv > 0) {
                    const float f_cos = rho * cosv / kPi, bsdf_pdf = cosv / kPi;
                    const float factor = scattering_direct(f_cos, bsdf_pdf, ls.pdf, ls.delta);
                    if (!std::isfinite(factor)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
                    const V3 from = offset_origin(sp.p, sp.err, sp.n, ls.wi);
                    V3 sd;
                    float tm;
                    if (ls.infinite) { sd = ls.wi; tm = INFINITY; }
                    else {
                        const V3 to = ls.delta ? ls.p : offset_origin(ls.p, ls.err, ls.n, from - ls.p);
                        sd = to - from;
                        tm = 1.0f - 1e-4f;
                    }
                    float tr[4];
                    uint32_t cr = 0;
                    ++out.shadow_rays;
                    if (transmittance(c, from, sd, tm, RP_NONE, wl, tr, cr))
                        L = L + throughput * factor * ls.le[0] * tr[0];
                }
            }
        }
        const V3 wi = hemisphere_dir(n, su(bw[0]), su(bw[1]));
        throughput = throughput * rho;
        if (!std::isfinite(throughput) || throughput < 0.0f) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
        ++out.scatterings;
        if (out.scatterings >= 3) {
            if (bw[2] >= 0xc0000000u) break;
            throughput = slip(SLIP_ROULETTE_UNCOMPENSATED) ? throughput : throughput / 0.75f;
        }
        const float cosine = dot(n, wi);
        if (!(cosine > 0.0f)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
        prev_pdf = cosine / kPi;
        escape_weight = demo ? (slip(SLIP_EMISSION_UNWEIGHTED) ? 1.0f : demo_bsdf_weight(cosine))
                             : power_heuristic(prev_pdf, 1.0f / (4.0f * kPi));
        prev_p = sp.p;
        prev_n = sp.n;
        o = offset_origin(sp.p, sp.err, sp.n, wi);
        d = wi;
        crossings = 0;
    }
    out.radiance = L;
    return out;
}
static void scalar_xyz(const Scene &s, const ScalarOut &p, float xyz[3]) {
    xyz[0] = xyz[1] = xyz[2] = 0.0f;
    if (p.status != RP_PATH_SUCCESS) return;
    if (s.d.transport == RP_TRANSPORT_SCALAR_DEMO) {
        demo_sensor(p.wavelength, p.radiance, xyz);
        return;
    }
    const auto &cie = cie_table();
    int i = int(std::floor(p.wavelength + 0.5f)) - 360;
    i = std::min(std::max(i, 0), 470);
    for (int ch = 0; ch < 3; ++ch) xyz[ch] = cie.xyz[i][ch] * (470.0f * p.radiance);
}

// Hero sensor: sum_k cie(lambda_k) * (L_k / pdf_k / 4)  (shaders/frame_sensor.h:32-53).
static void sensor_xyz(const PathOut &p, float xyz[3]) {
    xyz[0] = xyz[1] = xyz[2] = 0.0f;
    const auto &cie = cie_table();
    for (int k = 0; k < 4; ++k) {
        if (!(p.wl.pdf[k] > 0)) continue;
        int i = int(std::floor(p.wl.lambda[k] + 0.5f)) - 360;
        i = std::min(std::max(i, 0), 470);
        float w = (p.L[k] / p.wl.pdf[k]) / 4.0f;
        for (int ch = 0; ch < 3; ++ch) xyz[ch] = xyz[ch] + cie.xyz[i][ch] * w;
    }
}

} // namespace orc

// ================================================================ C API
using namespace orc;

extern "C" {

int or_init(const char *sobol_path, const char *cie_path) {
    return (load_sobol_tables(sobol_path) && load_cie_table(cie_path)) ? 0 : -1;
}

void or_set_roulette(int enabled) { g_roulette.store(enabled != 0); }
void or_plant_slip(int slip_id) { g_slip.store(slip_id); }

uint64_t or_sobol_index(uint32_t m, uint32_t x, uint32_t y, uint64_t sample) { return sobol_index(m, x, y, sample); }
uint32_t or_sobol_word(uint64_t index, uint32_t dim, uint32_t mode, uint32_t seed) {
    return sobol_word(index, dim, mode, seed);
}
uint32_t or_dimension_hash(uint32_t dim, uint32_t seed) { return sobol_dimension_hash(dim, seed); }
float or_sobol_uniform(uint32_t w) { return sobol_uniform(w); }
float or_camera_uniform(uint32_t w, uint32_t m, uint32_t p) { return sobol_camera_uniform(w, m, p); }
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    auto r = philox({ctr[0], ctr[1], ctr[2], ctr[3]}, {key[0], key[1]});
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}
float or_philox_u01(uint32_t w) { return philox_u01(w); }
void or_sample_wavelengths(float u, uint32_t mode, float lambda[4], float pdf[4]) {
    Wavelengths w = sample_wavelengths(u, mode);
    for (int k = 0; k < 4; ++k) {
        lambda[k] = w.lambda[k];
        pdf[k] = w.pdf[k];
    }
}
void or_cie_row(uint32_t i, float out[3]) {
    for (int c = 0; c < 3; ++c) out[c] = cie_table().xyz[i][c];
}
float or_spectrum_eval(const rp_spectrum *s, float lambda, const float *pool) { return spectrum_eval(*s, lambda, pool); }
void or_sincos_turns(float t, float *s, float *c) { sincos_turns(t, *s, *c); }
float or_exp(float x) { return exp_f(x); }
float or_log(float x) { return log_f(x); }

// BSDF probes for analytic pins (kind: rp_material_kind; roughness already alpha if remap==0).
void or_bsdf_eval(const rp_material *m, const float *spectra, float lambda, const float wo[3], const float wi[3],
                  float *f, float *pdf) {
    Scene s;
    s.spectra.assign(spectra ? spectra : (const float *)nullptr, spectra ? spectra + 471 : nullptr);
    Lane l = prepare_lane(s, *m, lambda);
    Eval e = eval_lane(l, {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]});
    *f = e.f;
    *pdf = e.pdf;
}
int or_bsdf_sample(const rp_material *m, const float *spectra, float lambda, const float wo[3], float uc, float u0,
                   float u1, float wi[3], float *pdf, int *delta) {
    Scene s;
    s.spectra.assign(spectra ? spectra : (const float *)nullptr, spectra ? spectra + 471 : nullptr);
    Lane l = prepare_lane(s, *m, lambda);
    Sampled r = sample_lane(l, {wo[0], wo[1], wo[2]}, uc, u0, u1);
    wi[0] = r.wi.x;
    wi[1] = r.wi.y;
    wi[2] = r.wi.z;
    *pdf = r.pdf;
    *delta = r.delta;
    return r.ok ? 1 : 0;
}
float or_fresnel_complex(float c, float eta, float k) { return fresnel_complex(c, eta, k); }
float or_fresnel_dielectric(float c, float eta) { return fresnel_dielectric(c, eta); }
float or_hg(float cos_t, float g) { return hg(cos_t, g); }
void or_hg_sample(const float wo[3], float g, float u0, float u1, float wi[3], float *pdf) {
    V3 r = hg_sample({wo[0], wo[1], wo[2]}, g, u0, u1, *pdf);
    wi[0] = r.x;
    wi[1] = r.y;
    wi[2] = r.z;
}
double or_conductor_dlogf(const rp_material *m, float lambda, const float wo[3], const float wi[3]) {
    Scene s;
    Lane l = prepare_lane(s, *m, lambda);
    return conductor_dlogf(l, *m, {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]});
}

int or_intersect_triangle(const float o[3], const float d[3], float tmax, const float p[9], float out[4]) {
    float t, b0, b1, b2;
    bool hit = intersect_triangle({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, tmax, {p[0], p[1], p[2]},
                                  {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, t, b0, b1, b2);
    out[0] = t;
    out[1] = b0;
    out[2] = b1;
    out[3] = b2;
    return hit ? 1 : 0;
}

void *or_scene_create(const rp_scene_desc *d) {
    auto *s = new Scene();
    s->d = *d;
    s->positions.assign(d->positions, d->positions + 3 * size_t(d->vertex_count));
    s->indices.assign(d->indices, d->indices + 3 * size_t(d->triangle_count));
    s->tri_object.assign(d->tri_object, d->tri_object + d->triangle_count);
    s->objects.assign(d->objects, d->objects + d->object_count);
    s->materials.assign(d->materials, d->materials + d->material_count);
    s->emissions.assign(d->emissions, d->emissions + d->emission_count);
    s->media.assign(d->media, d->media + d->medium_count);
    s->spots.assign(d->spots, d->spots + d->spot_count);
    s->spectra.assign(d->spectra, d->spectra + d->spectra_count);
    s->m = sobol_scale_exponent(d->camera.width, d->camera.height);
    compute_radius(*s);
    build_bvh(*s);
    build_lights(*s);
    return s;
}
void or_scene_destroy(void *s) { delete static_cast<Scene *>(s); }

uint32_t or_light_count(void *sp) { return uint32_t(static_cast<Scene *>(sp)->lights.size()); }
// Light table: kind, triangle / spot index, power-sampler pmf, alias
// threshold, 1/area, light-BVH leaf node.
void or_light_table(void *sp, uint32_t *kind, uint32_t *index, float *pmf, float *threshold, float *inv_area,
                    uint32_t *leaf) {
    auto *s = static_cast<Scene *>(sp);
    for (size_t i = 0; i < s->lights.size(); ++i) {
        kind[i] = s->lights[i].kind;
        index[i] = s->lights[i].index;
        pmf[i] = s->lights[i].pmf;
        threshold[i] = s->lights[i].threshold;
        inv_area[i] = s->lights[i].inv_area;
        leaf[i] = s->lights[i].leaf;
    }
}
uint32_t or_light_node_count(void *sp) { return uint32_t(static_cast<Scene *>(sp)->lnodes.size()); }
// The quad proposal of light `id` (a rectangle light) from p with receiver
// normal n (zero: none), as sample_light draws it: rectangle coordinates
// (u along ex, v along ey) and the solid-angle density.  Returns 1 when the
// spherical-rectangle proposal was used, 0 for area sampling, -1 when `id`
// is not a quad light.
int or_quad_sample(void *sp, uint32_t id, const float p[3], const float n[3], float u0, float u1, float uv[2],
                   float *pdf_omega) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (id >= s.lights.size() || s.lights[id].kind != 3) return -1;
    const Light &L = s.lights[id];
    const V3 r = {p[0], p[1], p[2]}, nn = {n[0], n[1], n[2]};
    float su = std::min(u0, 0x1.fffffep-1f), sv = std::min(u1, 0x1.fffffep-1f);
    uv[0] = su;
    uv[1] = sv;
    if (L.rect) {
        const RectFrame f = rect_frame(L, r);
        if (f.area > 1e-4f) {
            float dens = 1.0f;
            if (nn.x != 0.0f || nn.y != 0.0f || nn.z != 0.0f) {
                const Bilinear w = rect_cosines(f, L, nn);
                bilinear_warp(su, sv, w);
                dens = bilinear_density(su, sv, w);
            }
            rect_sample(f, su, sv, uv[0], uv[1]);
            *pdf_omega = dens / f.area;
            return 1;
        }
    }
    *pdf_omega = quad_area_pdf(L, r, L.p00 + uv[0] * L.ex + uv[1] * L.ey);
    return 0;
}
// The proposal's solid-angle density at rectangle coordinates uv (the
// emission-hit side of MIS).
float or_quad_pdf(void *sp, uint32_t id, const float p[3], const float n[3], const float uv[2]) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (id >= s.lights.size() || s.lights[id].kind != 3) return -1.0f;
    const Light &L = s.lights[id];
    const V3 q = L.p00 + uv[0] * L.ex + uv[1] * L.ey;
    return quad_pdf(L, {p[0], p[1], p[2]}, {n[0], n[1], n[2]}, uv[0], uv[1], q);
}
// The triangle proposal of light `id` (an emissive triangle) from p with
// receiver normal n: barycentrics (b1, b2) and the solid-angle density.
// Returns 1 for the spherical-triangle proposal, 0 for area sampling, -1
// when `id` is not a triangle light or the sample failed.
int or_tri_sample(void *sp, uint32_t id, const float p[3], const float n[3], float u0, float u1, float bary[2],
                  float *pdf_omega) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (id >= s.lights.size() || s.lights[id].kind != 0) return -1;
    const Light &L = s.lights[id];
    V3 a, b, c;
    s.tri(L.index, a, b, c);
    const V3 r = {p[0], p[1], p[2]}, nn = {n[0], n[1], n[2]};
    float su = std::min(u0, 0x1.fffffep-1f), sv = std::min(u1, 0x1.fffffep-1f);
    const float area = tri_solid_angle(a, b, c, r);
    if (spherical_tri(area)) {
        float dens = 1.0f;
        if (nn.x != 0.0f || nn.y != 0.0f || nn.z != 0.0f) {
            const Bilinear w = tri_cosines(a, b, c, r, nn);
            bilinear_warp(su, sv, w);
            dens = bilinear_density(su, sv, w);
        }
        if (!tri_sample(a, b, c, r, su, sv, bary[0], bary[1])) return -1;
        *pdf_omega = dens / area;
        return 1;
    }
    const float root = std::sqrt(su);
    bary[0] = 1.0f - root;
    bary[1] = sv * root;
    const V3 q = (1.0f - bary[0] - bary[1]) * a + bary[0] * b + bary[1] * c;
    *pdf_omega = tri_pdf(L, a, b, c, r, nn, q, normalize(cross(a - c, b - c)));
    return 0;
}
// The proposal's density at the triangle point with barycentrics (b1, b2).
float or_tri_pdf(void *sp, uint32_t id, const float p[3], const float n[3], const float bary[2]) {
    const Scene &s = *static_cast<Scene *>(sp);
    if (id >= s.lights.size() || s.lights[id].kind != 0) return -1.0f;
    const Light &L = s.lights[id];
    V3 a, b, c;
    s.tri(L.index, a, b, c);
    const V3 q = (1.0f - bary[0] - bary[1]) * a + bary[0] * b + bary[1] * c;
    return tri_pdf(L, a, b, c, {p[0], p[1], p[2]}, {n[0], n[1], n[2]}, q, normalize(cross(a - c, b - c)));
}
// A sequence of hero events from unit densities: before step i, the balance
// weight a / (a + b) of a sample drawn with densities pdf_i against the
// alternative other_i; then the cumulative update with (num_i, pdf_i).
// Returns the number of steps whose update succeeded.
uint32_t or_hero_sequence(uint32_t steps, const float *beta0, const float *num, const float *pdf, const float *other,
                          float *beta_out, float *weight_out) {
    float density[4] = {1, 1, 1, 1}, beta[4];
    for (int k = 0; k < 4; ++k) beta[k] = beta0[k];
    for (uint32_t i = 0; i < steps; ++i) {
        const float a = mixture(density, pdf + 4 * i), b = mixture(density, other + 4 * i);
        weight_out[i] = a / (a + b);
        if (!advance(density, beta, num + 4 * i, pdf + 4 * i)) return i;
        for (int k = 0; k < 4; ++k) beta_out[4 * i + k] = beta[k];
    }
    return steps;
}
// What the light tables were built from (one LightSamplingInput per light)
// and what they hold: alias links, and the light BVH's nodes.
void or_light_inputs(void *sp, double *power, double *bounds_power, float *low, float *high, float *dir,
                     float *cos_outer, float *cos_emission, uint8_t *two_sided, uint8_t *infinite) {
    const auto &in = static_cast<Scene *>(sp)->light_inputs;
    for (size_t i = 0; i < in.size(); ++i) {
        power[i] = in[i].power;
        bounds_power[i] = in[i].bounds_power;
        for (int a = 0; a < 3; ++a) {
            low[3 * i + a] = in[i].low[a];
            high[3 * i + a] = in[i].high[a];
            dir[3 * i + a] = in[i].dir[a];
        }
        cos_outer[i] = in[i].cos_outer;
        cos_emission[i] = in[i].cos_emission;
        two_sided[i] = in[i].two_sided ? 1 : 0;
        infinite[i] = in[i].infinite ? 1 : 0;
    }
}
void or_light_alias(void *sp, uint32_t *alias) {
    const auto &l = static_cast<Scene *>(sp)->lights;
    for (size_t i = 0; i < l.size(); ++i) alias[i] = l[i].alias;
}
void or_light_nodes(void *sp, float *low, float *high, float *dir, float *power, float *cos_outer, float *cos_emission,
                    uint32_t *child, uint32_t *parent, uint32_t *light, uint32_t *flags) {
    const auto &nodes = static_cast<Scene *>(sp)->lnodes;
    for (size_t k = 0; k < nodes.size(); ++k) {
        const LightNode &n = nodes[k];
        for (int a = 0; a < 3; ++a) {
            low[3 * k + a] = n.low[a];
            high[3 * k + a] = n.high[a];
            dir[3 * k + a] = n.dir[a];
        }
        power[k] = n.power;
        cos_outer[k] = n.cos_outer;
        cos_emission[k] = n.cos_emission;
        child[k] = n.child;
        parent[k] = n.parent;
        light[k] = n.light;
        flags[k] = n.flags;
    }
}
// Probes of the table helpers: the power proxy of a 471-sample table, the
// patch rectangle test of p = (p00, p10, p01, p11), the normal cone.
double or_spectrum_power_table(const float *table) {
    rp_spectrum sp{};
    sp.kind = RP_SPECTRUM_TABLE;
    sp.scale = 1;
    sp.table = 0;
    return spectrum_power_pool(sp, table);
}
int or_patch_rectangle(const double *p) {
    const D3 c[4] = {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, {p[9], p[10], p[11]}};
    return patch_rectangle(c) ? 1 : 0;
}
void or_normal_cone(const double *ns, uint32_t n, float *dir, float *cos_outer) {
    std::vector<D3> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = {ns[3 * i], ns[3 * i + 1], ns[3 * i + 2]};
    LightInput b;
    normal_cone(b, v.data(), int(n));
    for (int a = 0; a < 3; ++a) dir[a] = b.dir[a];
    *cos_outer = b.cos_outer;
}
// Light selection probes: the light chosen for (p, n, u) with its pmf, and
// the pmf of a given light.
// Test switch: `sampling` (an oracle scene of the same geometry, or NULL)
// drives the sampling decisions and densities of `sp` (Scene::sampling).
void or_scene_set_sampling(void *sp, void *sampling) {
    static_cast<Scene *>(sp)->sampling = static_cast<const Scene *>(sampling);
}

uint32_t or_light_select(void *sp, const float p[3], const float n[3], float u, float *pmf) {
    Selected sel = light_select(*static_cast<Scene *>(sp), {p[0], p[1], p[2]}, {n[0], n[1], n[2]}, u);
    *pmf = sel.pmf;
    return sel.light;
}
float or_light_pmf(void *sp, const float p[3], const float n[3], uint32_t id) {
    return light_pmf(*static_cast<Scene *>(sp), {p[0], p[1], p[2]}, {n[0], n[1], n[2]}, id);
}
// A full light sample from (p, n): light id (RP_NONE if none), point (or
// direction for the environment), pdf (selection pmf included), delta flag.
uint32_t or_light_sample(void *sp, const float p[3], const float n[3], float ul, float u0, float u1, float out[3],
                         float *pdf, int *delta) {
    auto *s = static_cast<Scene *>(sp);
    Wavelengths wl = sample_wavelengths(0.5f, 0);
    const V3 pp = {p[0], p[1], p[2]}, nn = {n[0], n[1], n[2]};
    const Selected sel = light_select(*s, pp, nn, ul);
    LightSample ls = sample_light(*s, pp, nn, wl, ul, u0, u1);
    *pdf = ls.ok ? ls.pdf : 0.0f;
    *delta = ls.delta ? 1 : 0;
    const V3 o = ls.infinite ? ls.wi : ls.p;
    out[0] = o.x;
    out[1] = o.y;
    out[2] = o.z;
    return ls.ok ? sel.light : RP_NONE;
}
// Solid-angle density of reaching emitter point q (geometric normal ng) of
// light `id` by light sampling from (p, n).
float or_area_light_pdf(void *sp, uint32_t id, const float p[3], const float n[3], const float q[3],
                        const float ng[3]) {
    return area_light_pdf(*static_cast<Scene *>(sp), id, {p[0], p[1], p[2]}, {n[0], n[1], n[2]}, {q[0], q[1], q[2]},
                          {ng[0], ng[1], ng[2]});
}
float or_atan2(float y, float x) { return atan2_f(y, x); }
float or_asin(float x) { return asin_f(x); }

// Closest hits for caller rays (rp_ray / rp_hit layouts).
void or_trace_closest(void *sp, const rp_ray *rays, uint64_t n, rp_hit *hits, uint32_t threads) {
    auto *s = static_cast<Scene *>(sp);
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint64_t> next{0};
    auto work = [&] {
        for (;;) {
            uint64_t b = next.fetch_add(1024);
            if (b >= n) break;
            uint64_t e = std::min(n, b + 1024);
            for (uint64_t i = b; i < e; ++i) {
                const rp_ray &r = rays[i];
                Hit h = closest(*s, {r.origin[0], r.origin[1], r.origin[2]},
                                {r.direction[0], r.direction[1], r.direction[2]}, r.tmax);
                hits[i].triangle = h.tri;
                hits[i].t = h.t;
                hits[i].b1 = h.b1;
                hits[i].b2 = h.b2;
                if (h.tri == RP_NONE) hits[i].t = INFINITY, hits[i].b1 = hits[i].b2 = 0;
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

// Per-path records for request (lane = pixel_offset + pixel_count * sample_offset).
int or_sample(void *sp, const rp_request *rq, rp_path_record *out, uint32_t threads) {
    auto *s = static_cast<Scene *>(sp);
    Ctx c{s, rq->seed, rq->randomization, rq->emergency_limit ? rq->emergency_limit : 1024u,
          rq->boundary_limit ? rq->boundary_limit : 4096u};
    uint64_t n = uint64_t(rq->pixel_count) * rq->sample_count;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint64_t> next{0};
    auto work = [&] {
        for (;;) {
            uint64_t b = next.fetch_add(256);
            if (b >= n) break;
            uint64_t e = std::min(n, b + 256);
            for (uint64_t lane = b; lane < e; ++lane) {
                uint32_t pixel = rq->pixel_begin + uint32_t(lane % rq->pixel_count);
                uint64_t sample = rq->sample_begin + lane / rq->pixel_count;
                PathOut p = trace_path(c, pixel, sample);
                rp_path_record &r = out[lane];
                std::memset(&r, 0, sizeof r);
                r.pixel = pixel;
                r.status = p.status;
                r.sample = sample;
                r.sobol_index = p.index;
                for (int k = 0; k < 4; ++k) {
                    r.wavelengths[k] = p.wl.lambda[k];
                    r.wavelength_pdfs[k] = p.wl.pdf[k];
                    r.radiance[k] = p.L[k];
                }
                if (p.status == RP_PATH_SUCCESS) sensor_xyz(p, r.xyz);
                r.primary_triangle = p.primary;
                r.vertices = p.vertices;
                r.camera_words[0] = p.w0;
                r.camera_words[1] = p.w1;
                r.filter_weight = 1.0f; // box pixel filter
                r.splat_begin = r.splat_count = 0;
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    return 0;
}

// Full frame: per-pixel mean XYZ of samples [sample_begin, sample_begin+spp),
// Neumaier-compensated double sums in increasing sample order (src/render.cpp:2326-2407).
} // extern "C"

template <typename Out>
static int render_frame_impl(void *sp, uint64_t seed, uint32_t randomization, uint64_t sample_begin, uint32_t spp,
                             Out *xyz, uint32_t threads, uint64_t *rays, uint32_t flags) {
    auto *s = static_cast<Scene *>(sp);
    const bool scalar = s->d.transport != RP_TRANSPORT_HERO;
    const bool independent = (flags & RP_FRAME_INDEPENDENT_SAMPLER) != 0;
    Ctx c{s, seed, randomization, 1024u, 4096u};
    uint32_t npx = s->d.camera.width * s->d.camera.height;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    std::atomic<uint64_t> ray_count{0};
    auto work = [&] {
        uint64_t local_rays = 0;
        for (;;) {
            uint32_t b = next.fetch_add(64);
            if (b >= npx) break;
            uint32_t e = std::min(npx, b + 64);
            for (uint32_t pixel = b; pixel < e; ++pixel) {
                double sum[3] = {0, 0, 0}, corr[3] = {0, 0, 0};
                for (uint32_t k = 0; k < spp; ++k) {
                    float v[3] = {0, 0, 0};
                    if (scalar) {
                        const ScalarOut p = trace_scalar(c, pixel, sample_begin + k, independent);
                        local_rays += p.ext_rays + p.shadow_rays;
                        scalar_xyz(*s, p, v);
                    } else {
                        PathOut p = trace_path(c, pixel, sample_begin + k);
                        local_rays += p.ext_rays + p.shadow_rays;
                        if (p.status == RP_PATH_SUCCESS) sensor_xyz(p, v);
                    }
                    for (int ch = 0; ch < 3; ++ch) {
                        double x = double(v[ch]);
                        double t = sum[ch] + x;
                        if (std::fabs(sum[ch]) >= std::fabs(x)) corr[ch] += (sum[ch] - t) + x;
                        else corr[ch] += (x - t) + sum[ch];
                        sum[ch] = t;
                    }
                }
                for (int ch = 0; ch < 3; ++ch) xyz[3 * size_t(pixel) + ch] = Out((sum[ch] + corr[ch]) / double(spp));
            }
        }
        ray_count += local_rays;
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (rays) *rays = ray_count.load();
    return 0;
}

extern "C" {

int or_render_frame_flags(void *sp, uint64_t seed, uint32_t randomization, uint64_t sample_begin, uint32_t spp,
                          float *xyz, uint32_t threads, uint64_t *rays, uint32_t flags) {
    return render_frame_impl(sp, seed, randomization, sample_begin, spp, xyz, threads, rays, flags);
}

// The film's per-pixel means in binary64 (no final rounding to binary32).
int or_render_frame_f64(void *sp, uint64_t seed, uint32_t randomization, uint64_t sample_begin, uint32_t spp,
                        double *xyz, uint32_t threads, uint64_t *rays, uint32_t flags) {
    return render_frame_impl(sp, seed, randomization, sample_begin, spp, xyz, threads, rays, flags);
}

int or_render_frame(void *sp, uint64_t seed, uint32_t randomization, uint64_t sample_begin, uint32_t spp,
                    float *xyz, uint32_t threads, uint64_t *rays) {
    return or_render_frame_flags(sp, seed, randomization, sample_begin, spp, xyz, threads, rays, 0);
}

// Scalar-transport records (rp_render_sample) for request rq; the sampler
// is rq->sampler_kind (Sobol or independent).
int or_sample_scalar(void *sp, const rp_request *rq, rp_render_sample *out, uint32_t threads) {
    auto *s = static_cast<Scene *>(sp);
    Ctx c{s, rq->seed, rq->randomization, rq->emergency_limit ? rq->emergency_limit : 1024u,
          rq->boundary_limit ? rq->boundary_limit : 4096u};
    const bool independent = rq->sampler_kind == RP_SAMPLER_INDEPENDENT;
    const uint64_t n = uint64_t(rq->pixel_count) * rq->sample_count;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint64_t> next{0};
    auto work = [&] {
        for (;;) {
            const uint64_t b = next.fetch_add(256);
            if (b >= n) break;
            const uint64_t e = std::min(n, b + 256);
            for (uint64_t lane = b; lane < e; ++lane) {
                const uint32_t pixel = rq->pixel_begin + uint32_t(lane % rq->pixel_count);
                const uint64_t sample = rq->sample_begin + lane / rq->pixel_count;
                const ScalarOut p = trace_scalar(c, pixel, sample, independent);
                rp_render_sample &r = out[lane];
                r.pixel = pixel;
                r.status = p.status;
                r.sample = sample;
                r.wavelength = p.wavelength;
                r.radiance = p.radiance;
                r.scatterings = p.scatterings;
                for (int k = 0; k < 4; ++k) r.random_words[k] = p.words[k];
                scalar_xyz(*s, p, r.xyz);
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    return 0;
}

// Reference gradient of sum_pixels dl_dxyz . XYZ (forward-mode, double accumulation).
// Parameter layout as rp_backward: binned spectra consecutive with bin 0 first.
int or_gradient(void *sp, uint64_t seed, uint32_t randomization, uint64_t sample_begin, uint32_t spp,
                const rp_param *params, uint32_t count, const float *dl_dxyz, double *grad_out, uint32_t threads) {
    auto *s = static_cast<Scene *>(sp);
    GradSpec g;
    g.nparams = count;
    g.refl_base.assign(s->materials.size() + 1, -1);
    g.rough.assign(s->materials.size() + 1, -1);
    g.refl_bins.assign(s->materials.size() + 1, 0);
    g.sigma_base.assign(s->media.size() + 1, -1);
    g.sigma_bins.assign(s->media.size() + 1, 0);
    for (uint32_t i = 0; i < count; ++i) {
        const rp_param &p = params[i];
        if (p.kind == RP_PARAM_ROUGHNESS) g.rough[p.target] = int32_t(i);
        else if (p.kind == RP_PARAM_REFLECTANCE_BIN && p.bin == 0) g.refl_base[p.target] = int32_t(i), g.refl_bins[p.target] = p.bins;
        else if (p.kind == RP_PARAM_SIGMA_S_BIN && p.bin == 0) {
            g.sigma_base[p.target] = int32_t(i), g.sigma_bins[p.target] = p.bins;
            g.sigma_medium = p.target;
        }
    }
    Ctx c{s, seed, randomization, 1024u, 4096u};
    const uint32_t npx = s->d.camera.width * s->d.camera.height;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::vector<double>> partial(threads, std::vector<double>(count, 0.0));
    std::atomic<uint32_t> next{0};
    auto work = [&](uint32_t tid) {
        ForwardGrad fg;
        fg.g = &g;
        fg.grad = &partial[tid];
        for (;;) {
            uint32_t b = next.fetch_add(16);
            if (b >= npx) break;
            for (uint32_t pixel = b; pixel < std::min(npx, b + 16); ++pixel) {
                for (uint32_t k = 0; k < spp; ++k) {
                    // dLoss/dL_k for this path needs its wavelengths: trace once to get them
                    uint64_t sample = sample_begin + k;
                    uint32_t px = pixel % s->d.camera.width, py = pixel / s->d.camera.width;
                    uint64_t index = sobol_index(s->m, px, py, sample);
                    uint32_t w2 = sobol_word(index, 2, randomization, uint32_t(seed));
                    Wavelengths wl = sample_wavelengths(sobol_uniform(w2), s->d.hero_wavelengths);
                    for (int l = 0; l < 4; ++l) {
                        fg.a[l] = 0.0;
                        if (wl.pdf[l] > 0) {
                            int i = std::min(std::max(int(std::floor(wl.lambda[l] + 0.5f)) - 360, 0), 470);
                            const auto &cie = cie_table().xyz[i];
                            double d = double(dl_dxyz[3 * size_t(pixel)]) * cie[0] + double(dl_dxyz[3 * size_t(pixel) + 1]) * cie[1] +
                                       double(dl_dxyz[3 * size_t(pixel) + 2]) * cie[2];
                            fg.a[l] = d / double(wl.pdf[l]) / (4.0 * spp);
                        }
                    }
                    std::vector<double> before = partial[tid];
                    PathOut po = trace_path_t(c, pixel, sample, fg);
                    if (po.status != RP_PATH_SUCCESS) partial[tid] = before; // failed paths contribute nothing
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &t : pool) t.join();
    for (uint32_t p = 0; p < count; ++p) {
        double acc = 0;
        for (uint32_t t = 0; t < threads; ++t) acc += partial[t][p];
        grad_out[p] = acc;
    }
    return 0;
}

} // extern "C"
