// Device-side transport for gfx950: math, sampling, BVH4 traversal, BSDFs,
// lights, media and the hero-wavelength path event.  Everything here is
// binary32 with no FMA contraction (the library is built with
// -ffp-contract=off) and uses only the approximations defined in this file
// for transcendentals, so per-path results are reproducible bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rp_api.h"

#define RP_DEV __device__ __forceinline__
#define RP_NOINLINE __device__ __attribute__((noinline))

namespace rpd {

// ============================================================== vectors
struct F3 {
    float x, y, z;
};
RP_DEV F3 f3(float x, float y, float z) { return F3{x, y, z}; }
RP_DEV F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RP_DEV F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RP_DEV F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
RP_DEV F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
RP_DEV F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RP_DEV F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RP_DEV float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RP_DEV F3 vabs(F3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
RP_DEV float comp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// std::min / std::max semantics (returns the first argument on ties / NaN).
RP_DEV float omin(float a, float b) { return (b < a) ? b : a; }
RP_DEV float omax(float a, float b) { return (a < b) ? b : a; }
// std::clamp semantics.
RP_DEV float oclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
RP_DEV float dop(float a, float b, float c, float d) {
    float cd = c * d;
    float err = __builtin_fmaf(-c, d, cd);
    float r = __builtin_fmaf(a, b, -cd);
    return r + err;
}
RP_DEV F3 cross(F3 a, F3 b) { return {dop(a.y, b.z, a.z, b.y), dop(a.z, b.x, a.x, b.z), dop(a.x, b.y, a.y, b.x)}; }
// Plain cross product (the light-sampling geometry uses this form).
RP_DEV F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RP_DEV float length(F3 a) { return __builtin_sqrtf(dot(a, a)); }
RP_DEV F3 normalize(F3 a) { return a / length(a); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv4Pi = 0.07957747154594766788f;
RP_DEV float gamma_n(int n) { return (float(n) * 0x1p-24f) / (1.0f - float(n) * 0x1p-24f); }

// ======================================================= transcendentals
// sin/cos of 2*pi*t: exact quadrant reduction in turns, Taylor on |a|<=pi/4.
RP_DEV void sincos_turns(float t, float &s, float &c) {
    float x = t * 4.0f;
    float q = floorf(x + 0.5f);
    float r = (x - q) * 0.25f;
    float a = r * 6.28318530717958647692f;
    float a2 = a * a;
    float sp = a + a * a2 * (-1.0f / 6.0f + a2 * (1.0f / 120.0f + a2 * (-1.0f / 5040.0f + a2 * (1.0f / 362880.0f))));
    float cp = 1.0f + a2 * (-0.5f + a2 * (1.0f / 24.0f + a2 * (-1.0f / 720.0f + a2 * (1.0f / 40320.0f))));
    int qi = int(q) & 3;
    if (qi == 0) { s = sp; c = cp; }
    else if (qi == 1) { s = cp; c = -sp; }
    else if (qi == 2) { s = -sp; c = -cp; }
    else { s = -cp; c = sp; }
}
// exp: Cody-Waite two-part ln2 reduction, degree-7 Taylor.
RP_DEV float exp_f(float x) {
    if (x != x) return x;
    if (!(x > -104.0f)) return 0.0f;
    if (x > 88.72f) return __builtin_inff();
    float k = floorf(x * 1.44269504088896341f + 0.5f);
    float r = (x - k * 0.693145751953125f) - k * 1.42860682030941723212e-6f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (1.0f / 6.0f + r * (1.0f / 24.0f + r * (1.0f / 120.0f + r * (1.0f / 720.0f + r * (1.0f / 5040.0f)))))));
    return ldexpf(p, int(k));
}
// log: frexp to [sqrt(1/2), sqrt(2)), atanh series in (m-1)/(m+1).
RP_DEV float log_f(float x) {
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e;
    float m = frexpf(x, &e);
    if (m < 0.70710678118654752f) { m = m * 2.0f; e -= 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float lm = s * (2.0f + s2 * (2.0f / 3.0f + s2 * (2.0f / 5.0f + s2 * (2.0f / 7.0f + s2 * (2.0f / 9.0f)))));
    float ef = float(e);
    return ef * 0.693145751953125f + (lm + ef * 1.42860682030941723212e-6f);
}
// sin and cos of an angle in radians through the turn reduction.
RP_DEV float sin_f(float x) {
    float s, c;
    sincos_turns(x * 0.159154943091895336f, s, c);
    return s;
}
RP_DEV float cos_f(float x) {
    float s, c;
    sincos_turns(x * 0.159154943091895336f, s, c);
    return c;
}
// atan: 1/|x| for |x| > 1 (pi/2 minus), then pi/4 + atan((z-1)/(z+1)) above
// tan(pi/8); the Maclaurin series to t^17 on |t| <= tan(pi/8).
RP_DEV float atan_f(float x) {
    const float ax = fabsf(x);
    const bool inv = ax > 1.0f;
    const float z = inv ? 1.0f / ax : ax;
    const bool shift = z > 0.414213562373095049f;
    const float t = shift ? (z - 1.0f) / (z + 1.0f) : z;
    const float t2 = t * t;
    const float p = t + t * t2 * (-1.0f / 3.0f + t2 * (1.0f / 5.0f + t2 * (-1.0f / 7.0f + t2 * (1.0f / 9.0f +
                    t2 * (-1.0f / 11.0f + t2 * (1.0f / 13.0f + t2 * (-1.0f / 15.0f + t2 * (1.0f / 17.0f))))))));
    float r = shift ? 0.785398163397448310f + p : p;
    r = inv ? 1.57079632679489662f - r : r;
    return copysignf(r, x);
}
RP_DEV float atan2_f(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) return signbit(x) ? copysignf(3.14159265358979324f, y) : copysignf(0.0f, y);
    float r = ay <= ax ? atan_f(ay / ax) : 1.57079632679489662f - atan_f(ax / ay);
    if (signbit(x)) r = 3.14159265358979324f - r;
    return copysignf(r, y);
}
RP_DEV float asin_f(float x) { return atan2_f(x, __builtin_sqrtf(omax(0.0f, (1.0f - x) * (1.0f + x)))); }

// Light-sampling context normal: a unit normal through a 2 x 16-bit
// octahedral code (snorm, round half up) and back, so that a vertex's light
// sampling and the next vertex's MIS weight use the same normal bits; the
// code also fits the path state's spare word.  kNoNormal (media, camera)
// decodes to zero.
constexpr uint32_t kNoNormal = 0xffffffffu;
RP_DEV float sign_nz(float v) { return v < 0.0f ? -1.0f : 1.0f; }
RP_DEV uint32_t oct_encode(F3 n) {
    const float l1 = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float px = n.x / l1, py = n.y / l1;
    if (n.z < 0.0f) {
        const float qx = (1.0f - fabsf(py)) * sign_nz(px), qy = (1.0f - fabsf(px)) * sign_nz(py);
        px = qx;
        py = qy;
    }
    const int32_t ix = int32_t(floorf(oclamp(px, -1.0f, 1.0f) * 32767.0f + 0.5f));
    const int32_t iy = int32_t(floorf(oclamp(py, -1.0f, 1.0f) * 32767.0f + 0.5f));
    return (uint32_t(ix) & 0xffffu) | (uint32_t(iy) << 16);
}
RP_DEV F3 oct_decode(uint32_t c) {
    if (c == kNoNormal) return F3{0, 0, 0};
    float px = float(int16_t(c & 0xffffu)) / 32767.0f, py = float(int16_t(c >> 16)) / 32767.0f;
    const float z = 1.0f - fabsf(px) - fabsf(py);
    if (z < 0.0f) {
        const float qx = (1.0f - fabsf(py)) * sign_nz(px), qy = (1.0f - fabsf(px)) * sign_nz(py);
        px = qx;
        py = qy;
    }
    const F3 v = {px, py, z};
    return v / __builtin_sqrtf(dot(v, v));
}

RP_DEV float next_up(float v) {
    if (__builtin_isinf(v) && v > 0) return v;
    if (v == -0.0f) v = 0.0f;
    uint32_t u = __float_as_uint(v);
    if (v >= 0) ++u; else --u;
    return __uint_as_float(u);
}
RP_DEV float next_down(float v) {
    if (__builtin_isinf(v) && v < 0) return v;
    if (v == 0.0f) v = -0.0f;
    uint32_t u = __float_as_uint(v);
    if (v > 0) --u; else ++u;
    return __uint_as_float(u);
}

// ======================================================== scene on device
// One light (128 B).  kind 0 triangle, 1 spot, 2 environment, 3 quad (a
// rectangle p00 + u ex + v ey made of two triangles); index: the leaf-order
// position of the (first) triangle or the spot index; leaf: the light-BVH
// node holding the light; pmf / threshold / alias: the power sampler's alias
// table; rect: the quad passes the binary32 rectangle test; ux, uy, uz, lx,
// ly: the quad's unit frame and edge lengths.
struct LightEntry {
    uint32_t kind, index, leaf, rect;
    float pmf, threshold;
    uint32_t alias;
    float inv_area;
    float p00[3], ex[3], ey[3];
    float ux[3], uy[3], uz[3];
    float lx, ly;
    uint32_t pad[2];
};
// Light-BVH node (64 B): the bounding sphere of its box (centre, squared
// radius, radius), direction cone (axis, cosine and sine of its half-angle),
// emission cutoff, power normalised by the
// tree's total; the first child of interior node i is i + 1; flags bit 0 =
// two-sided.
struct LightNode {
    float center[3], r2, radius;
    float dir[3];
    float cos_outer, sin_outer, cos_emission, power;
    uint32_t child, parent, light, flags;
};

struct Scene {
    const float4 *tris;     // 3 per triangle in BVH leaf order: (v0, scene id) (v1, object) (v2, light)
    const uint32_t *tri_material; // material index per triangle, BVH leaf order (sort keys)
    const uint4 *nodes;     // 4 per 64-byte node
    uint32_t node_count, tri_count;
    const rp_object *objects;
    const rp_material *materials;
    const rp_emission *emissions;
    const rp_medium *media;
    const rp_spot_light *spots;
    const float *spectra;
    const LightEntry *lights;
    uint32_t light_count;
    uint32_t has_environment;
    const LightNode *light_nodes;
    uint32_t light_node_count;
    uint32_t light_sampler; // rp_light_sampler
    uint32_t env_light;     // the environment's light id (RP_NONE: none)
    const float4 *cie;          // 471 rows (x, y, z, 0)
    const uint32_t *sobol_dir;  // transposed directions: [52][1024]
    const uint64_t *sobol_intervals, *sobol_inverse;
    rp_camera camera;
    rp_spectrum environment;
    uint32_t max_depth, hero_mode, m;
    uint32_t has_interfaces, has_media;
    uint32_t transport; // rp_transport
    float env_pdf;
};

// =============================================================== sampling
RP_DEV uint64_t sobol_index(const Scene &S, uint32_t x, uint32_t y, uint64_t sample) {
    uint32_t m = S.m;
    if (m == 0) return sample;
    uint32_t row = (m - 1) * 52;
    uint64_t delta = 0;
    uint64_t result = sample << (2 * m);
    for (uint64_t s = sample; s; s &= s - 1) delta ^= S.sobol_intervals[row + __builtin_ctzll(s)];
    uint64_t bits = ((uint64_t(x) << m) | y) ^ delta;
    for (; bits; bits &= bits - 1) result ^= S.sobol_inverse[row + __builtin_ctzll(bits)];
    return result;
}
RP_DEV uint32_t dimension_hash(uint32_t dimension, uint32_t seed_bits) {
    const uint64_t mm = 0xc6a4a7935bd1e995ull;
    uint64_t k = ((uint64_t(seed_bits) << 32) | dimension) * mm;
    k ^= k >> 47;
    k *= mm;
    uint64_t h = (8 * mm) ^ k;
    h *= mm;
    h ^= h >> 47;
    h *= mm;
    h ^= h >> 47;
    return uint32_t(h);
}
RP_DEV uint64_t mix64(uint64_t v) {
    v ^= v >> 31;
    v *= 0x7fb5d329728ea185ull;
    v ^= v >> 27;
    v *= 0x81dadef4bc2dd44dull;
    return v ^ (v >> 33);
}
RP_DEV uint32_t scramble(uint32_t v, uint32_t dimension, uint32_t mode, uint32_t seed_bits) {
    if (mode == RP_RANDOMIZE_NONE) return v;
    uint32_t hash = dimension_hash(dimension, seed_bits);
    if (mode == RP_RANDOMIZE_PERMUTE) return v ^ hash;
    if (mode == RP_RANDOMIZE_FASTOWEN) {
        v = __builtin_bitreverse32(v);
        v ^= v * 0x3d20adeau;
        v += hash;
        v *= (hash >> 16) | 1u;
        v ^= v * 0x05526c56u;
        v ^= v * 0x53a22864u;
        return __builtin_bitreverse32(v);
    }
    v ^= (hash & 1u) << 31;
    for (uint32_t b = 1; b < 32; ++b) {
        uint32_t prefix = v & (0xffffffffu << (32 - b));
        if (uint32_t(mix64(uint64_t(prefix ^ hash))) & (1u << b)) v ^= 1u << (31 - b);
    }
    return v;
}
// The unscrambled words of dimension blocks a and b (4 dimensions each)
// from one pass over the index bits: both blocks' direction loads are issued
// together, early in the event.
RP_DEV void sobol_raw2(const Scene &S, uint64_t index, uint32_t a, uint32_t b, uint4 &va, uint4 &vb) {
    va = make_uint4(0, 0, 0, 0);
    vb = make_uint4(0, 0, 0, 0);
    const uint4 *dir4 = reinterpret_cast<const uint4 *>(S.sobol_dir);
    for (uint64_t i = index; i; i &= i - 1) {
        const uint32_t row = uint32_t(__builtin_ctzll(i)) * 1024u;
        const uint4 x = dir4[(row + 4 * a) >> 2], y = dir4[(row + 4 * b) >> 2];
        va.x ^= x.x, va.y ^= x.y, va.z ^= x.z, va.w ^= x.w;
        vb.x ^= y.x, vb.y ^= y.y, vb.z ^= y.z, vb.w ^= y.w;
    }
}
// Scrambles the raw words of dimension block `block`.
RP_DEV uint4 sobol_scramble4(uint4 v, uint32_t block, uint32_t mode, uint32_t seed_bits) {
    const uint32_t d = 4 * block;
    return make_uint4(scramble(v.x, d, mode, seed_bits), scramble(v.y, d + 1, mode, seed_bits),
                      scramble(v.z, d + 2, mode, seed_bits), scramble(v.w, d + 3, mode, seed_bits));
}
// Four consecutive dimensions 4*block..4*block+3 with one 16-byte load per index bit.
RP_DEV uint4 sobol_block(const Scene &S, uint64_t index, uint32_t block, uint32_t mode, uint32_t seed_bits) {
    uint32_t d = 4 * block;
    uint4 v = make_uint4(0, 0, 0, 0);
    const uint4 *dir4 = reinterpret_cast<const uint4 *>(S.sobol_dir);
    for (uint64_t i = index; i; i &= i - 1) {
        uint4 w = dir4[(uint32_t(__builtin_ctzll(i)) * 1024u + d) >> 2];
        v.x ^= w.x;
        v.y ^= w.y;
        v.z ^= w.z;
        v.w ^= w.w;
    }
    return make_uint4(scramble(v.x, d, mode, seed_bits), scramble(v.y, d + 1, mode, seed_bits),
                      scramble(v.z, d + 2, mode, seed_bits), scramble(v.w, d + 3, mode, seed_bits));
}
RP_DEV float sobol_uniform(uint32_t word) {
    float u = float(word) * 0x1p-32f;
    return u < 0x1.fffffep-1f ? u : 0x1.fffffep-1f;
}
RP_DEV float camera_uniform(uint32_t word, uint32_t m, uint32_t pixel) {
    float local = sobol_uniform(word) * float(uint32_t(1) << m) - float(pixel);
    if (local < 0.0f) return 0.0f;
    return local < 0x1.fffffep-1f ? local : 0x1.fffffep-1f;
}
RP_DEV uint4 philox(uint4 c, uint2 k) {
    const uint32_t m0 = 0xd2511f53u, m1 = 0xcd9e8d57u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(m0, c.x), lo0 = m0 * c.x;
        uint32_t hi1 = __umulhi(m1, c.z), lo1 = m1 * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9e3779b9u;
        k.y += 0xbb67ae85u;
    }
    return c;
}
RP_DEV float philox_u01(uint32_t w) { return float(w >> 8) * 0x1p-24f; }

struct Wavelengths {
    float lambda[4], pdf[4];
};
RP_DEV Wavelengths sample_wavelengths(float u, uint32_t mode) {
    Wavelengths w;
    if (!(u >= 0 && u < 1)) {
        for (int k = 0; k < 4; ++k) w.lambda[k] = w.pdf[k] = __builtin_nanf("");
        return w;
    }
    if (mode == 0) {
        for (int k = 0; k < 4; ++k) {
            float s = u + float(k) * 0.25f;
            if (s >= 1.0f) s -= 1.0f;
            float l = 360.0f + 470.0f * s;
            if (l >= 830.0f) l = __uint_as_float(__float_as_uint(830.0f) - 1);
            w.lambda[k] = l;
            w.pdf[k] = 1.0f / 470.0f;
        }
        return w;
    }
    for (int k = 0; k < 4; ++k) {
        uint32_t coord = uint32_t(u * 16777216.0f);
        float s = float((coord + uint32_t(k) * 4194304u) & 0x00ffffffu) * 0x1p-24f;
        float a = 0.85691062f - 1.82750197f * s;
        float at = 0.5f * log_f((1.0f + a) / (1.0f - a));
        float l = 538.0f - 138.888889f * at;
        l = omin(omax(l, 360.0f), 830.0f);
        float e = exp_f(0.0072f * (l - 538.0f));
        float ch = 0.5f * (e + 1.0f / e);
        w.lambda[k] = l;
        w.pdf[k] = 0.0039398042f / (ch * ch);
    }
    return w;
}

RP_DEV float spectrum_eval(const rp_spectrum &s, float l, const float *pool) {
    float v;
    switch (s.kind) {
    case RP_SPECTRUM_CONSTANT: v = s.c[0]; break;
    case RP_SPECTRUM_SIGMOID: {
        float x = __builtin_fmaf(__builtin_fmaf(s.c[0], l, s.c[1]), l, s.c[2]);
        if (__builtin_isinf(x)) v = x > 0 ? 1.0f : 0.0f;
        else v = 0.5f + x / (2.0f * __builtin_sqrtf(1.0f + x * x));
        break;
    }
    case RP_SPECTRUM_TABLE: {
        int i = int(floorf(l + 0.5f)) - 360;
        i = i < 0 ? 0 : (i > 470 ? 470 : i);
        v = pool[s.table + uint32_t(i)];
        break;
    }
    case RP_SPECTRUM_CAUCHY: {
        float um = l * 0.001f;
        v = s.c[0] + s.c[1] / (um * um);
        break;
    }
    case RP_SPECTRUM_LINEAR: {
        float t = (l - 360.0f) / 470.0f;
        v = (1.0f - t) * s.c[0] + t * s.c[1];
        break;
    }
    case RP_SPECTRUM_KNOTS8: {
        // piecewise linear through the 8 pool values at 360, 420, 470, 520,
        // 570, 620, 700 and 830 nm (x + (y - x) t between neighbours)
        const float *r = pool + s.table;
        const float knot[8] = {360.0f, 420.0f, 470.0f, 520.0f, 570.0f, 620.0f, 700.0f, 830.0f};
        v = r[7];
        for (int i = 1; i < 8; ++i)
            if (l <= knot[i]) {
                const float t = (l - knot[i - 1]) / (knot[i] - knot[i - 1]);
                v = r[i - 1] + (r[i] - r[i - 1]) * t;
                break;
            }
        break;
    }
    default: v = __builtin_nanf("");
    }
    return v * s.scale;
}

// ================================================================ geometry
struct Hit {
    uint32_t prim;   // leaf-order position, 0xffffffff = miss
    uint32_t id;     // scene triangle id
    float t, b0, b1, b2;
};

RP_DEV bool intersect_triangle(F3 o, F3 d, float tmax, F3 p0, F3 p1, F3 p2, float &t_out, float &b0o, float &b1o,
                               float &b2o) {
    F3 p0t = p0 - o, p1t = p1 - o, p2t = p2 - o;
    F3 ad = vabs(d);
    int kz = ad.x > ad.y ? (ad.x > ad.z ? 0 : 2) : (ad.y > ad.z ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    F3 dp = {comp(d, kx), comp(d, ky), comp(d, kz)};
    p0t = {comp(p0t, kx), comp(p0t, ky), comp(p0t, kz)};
    p1t = {comp(p1t, kx), comp(p1t, ky), comp(p1t, kz)};
    p2t = {comp(p2t, kx), comp(p2t, ky), comp(p2t, kz)};
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
    float max_zt = omax(omax(fabsf(p0t.z), fabsf(p1t.z)), fabsf(p2t.z));
    float delta_z = gamma_n(3) * max_zt;
    float max_xt = omax(omax(fabsf(p0t.x), fabsf(p1t.x)), fabsf(p2t.x));
    float max_yt = omax(omax(fabsf(p0t.y), fabsf(p1t.y)), fabsf(p2t.y));
    float delta_x = gamma_n(5) * (max_xt + max_zt);
    float delta_y = gamma_n(5) * (max_yt + max_zt);
    float delta_e = 2.0f * (gamma_n(2) * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
    float max_e = omax(omax(fabsf(e0), fabsf(e1)), fabsf(e2));
    float delta_t = 3.0f * (gamma_n(3) * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * fabsf(inv);
    if (t <= delta_t) return false;
    t_out = t;
    b0o = b0;
    b1o = b1;
    b2o = b2;
    return true;
}

// ------------------------------------------------------ BVH4 traversal
// Node layout (64 B = 4 x uint4):
//   q0: origin.xyz, exponent bytes (x, y, z) | child_count << 24
//   q1: qlo.x[4], qlo.y[4], qlo.z[4], qhi.x[4]  (one byte per child)
//   q2: qhi.y[4], qhi.z[4], child[0], child[1]
//   q3: child[2], child[3], count[4], pad
// Leaf child word: 0x80000000 | count << 27 | first (leaf-order triangle).
constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kEmpty = 0xffffffffu;
// Per-lane LDS entries of the traversal stack (node + entry distance, 8 B
// each).  LDS per 128-lane block = 1 KiB per entry, so this sets the
// occupancy ceiling of the traversal kernels; deeper entries spill to scratch.
// Measured on MI355X (1080p terrain wavefront frame): 4 entries 47.5 ms,
// 8 entries 46.3 ms, 12 entries 47.5 ms, 16 entries 49.8 ms.
#ifndef RP_SHORT_STACK
#define RP_SHORT_STACK 8
#endif
constexpr int kShortStack = RP_SHORT_STACK;

// Whole-path kernels (megakernels, latency tails, the persistent batch) inline
// the path loop, traversal and light sampling together.  Left to itself the
// compiler gives them 256 VGPRs plus accumulation registers, one wave per
// SIMD; two waves per SIMD (at most 256 registers in all) keep twice the
// memory requests in flight.
#ifndef RP_PATH_WAVES
#define RP_PATH_WAVES 2
#endif
#define RP_PATH_ATTR __attribute__((amdgpu_waves_per_eu(RP_PATH_WAVES)))
constexpr int kMaxStack = 64;                      // the BVH builder's bound is checked against this
constexpr int kSpillStack = kMaxStack - kShortStack; // per-lane overflow (scratch)
static_assert(kShortStack >= 1 && kShortStack <= kMaxStack, "short stack size");

// The short stack is addressed through the LDS address space: through a
// generic pointer the compiler emits flat loads and stores, which wait on
// (and are waited for with) every outstanding global load of the lane.
typedef __attribute__((address_space(3))) uint32_t LdsU32;
typedef __attribute__((address_space(3))) float LdsF32;

// Per-lane overflow entries, a private array of their own: kept apart from
// the Stack's scalar fields so that those stay in registers (a struct that
// holds a dynamically indexed array lives whole in scratch, stack pointer
// included, and every push and pop then waits on scratch).
struct SpillStack {
    uint32_t node[kSpillStack];
    float t[kSpillStack];
};
struct Stack {
    LdsU32 *lds_node; // this lane's column of [kShortStack][block]
    LdsF32 *lds_t;
    int stride;       // block size
    int sp;
    SpillStack *spill; // set where the Stack is declared (RP_STACK)
    // node, t: the block's __shared__ arrays of kShortStack * block entries.
    RP_DEV void init(uint32_t *node, float *t, int block) {
        lds_node = (LdsU32 *)(node + threadIdx.x);
        lds_t = (LdsF32 *)(t + threadIdx.x);
        stride = block;
        sp = 0;
    }
    RP_DEV void push(uint32_t n, float t) {
        if (sp < kShortStack) {
            lds_node[sp * stride] = n;
            lds_t[sp * stride] = t;
        } else {
            spill->node[sp - kShortStack] = n;
            spill->t[sp - kShortStack] = t;
        }
        ++sp;
    }
    RP_DEV void pop(uint32_t &n, float &t) {
        --sp;
        if (sp < kShortStack) {
            n = lds_node[sp * stride];
            t = lds_t[sp * stride];
        } else {
            n = spill->node[sp - kShortStack];
            t = spill->t[sp - kShortStack];
        }
    }
};
// Declares traversal stack `name` with its overflow array.
#define RP_STACK(name)                                                                                                 \
    ::rpd::Stack name;                                                                                                 \
    ::rpd::SpillStack name##_spill;                                                                                    \
    name.spill = &name##_spill

RP_DEV float byte_of(uint32_t w, int i) { return float((w >> (8 * i)) & 0xffu); }

// Reciprocal direction component for the slab test.  An exact zero gives
// +-2^100 instead of an infinity: q * A + B with A infinite is NaN, and an
// axis without a finite scale would cull nothing (a vertical ray over a
// heightfield would visit every node under it).  With 2^100 a ray parallel
// to the slab gets an interval of at least +-(pad + 2^-20 w) * 2^100 around
// 0 when its origin lies inside (pad >= 1e-30 is the builder's padding, w
// the child's extent on that axis), beyond any hit distance unless the
// child is thinner than ~1e-18 on that axis, whose geometry such a ray can
// only meet in its own plane, where the triangle test reports no hit; when
// the origin lies outside at a gap g, the interval starts near g * 2^100,
// which culls the box unless g is below ~1e-24 (then the box costs a visit,
// never a hit).  A and B stay finite for scenes under ~1e8 units.
RP_DEV float slab_inv(float d) { return d == 0.0f ? copysignf(0x1p100f, d) : 1.0f / d; }

// Tests the 4 children of a node; returns hit mask and entry distances.
//
// A child's plane on axis x sits at P = ox + q * sx (8-bit q, power-of-two
// sx).  Along the ray its distance is (P - o.x) / d.x = q * A + B with
// A = sx * inv.x (exact: sx is a power of two) and B = (ox - o.x) * inv.x,
// so each plane costs one byte conversion and one fused multiply-add; the
// direction's sign picks which of the two bounds is the near plane, so the
// slab needs no min/max.  The computed distances differ from the exact ones
// by at most 2^-24 * (510 |A| + 5.1 |B|) (the rounding of B, the reciprocal,
// the fused sum, and the widened offsets below); E = 2^-20 * (255 |A| +
// 2 |B|) exceeds that six-fold, and the near planes use B - E, the far
// planes B + E.
// The slab interval so computed contains the exact one, so no box that the
// ray meets is culled.  The builder pads each child box by 2^-20 of its
// magnitude before quantising, which keeps the exact P (not only its
// rounded value) outside the child's geometry.  `inv` comes from slab_inv
// (finite for a zero component); where A or B overflows on a tiny nonzero
// component, the distances are NaN, which min/max ignore, or infinite on
// the side the exact values lie.
RP_DEV uint32_t test_children(const uint4 *nodes, uint32_t node, F3 o, F3 inv, float tmax, float tn[4], uint32_t child[4]) {
    const uint4 q0 = nodes[4 * node + 0];
    const uint4 q1 = nodes[4 * node + 1];
    const uint4 q2 = nodes[4 * node + 2];
    const uint4 q3 = nodes[4 * node + 3];
    const float sx = __uint_as_float((q0.w & 0xffu) << 23);
    const float sy = __uint_as_float(((q0.w >> 8) & 0xffu) << 23);
    const float sz = __uint_as_float(((q0.w >> 16) & 0xffu) << 23);
    child[0] = q2.z;
    child[1] = q2.w;
    child[2] = q3.x;
    child[3] = q3.y;
    const float ax = sx * inv.x, ay = sy * inv.y, az = sz * inv.z;
    const float bx = (__uint_as_float(q0.x) - o.x) * inv.x;
    const float by = (__uint_as_float(q0.y) - o.y) * inv.y;
    const float bz = (__uint_as_float(q0.z) - o.z) * inv.z;
    const float ex = __builtin_fmaf(fabsf(bx), 0x1p-19f, fabsf(ax) * (255.0f * 0x1p-20f));
    const float ey = __builtin_fmaf(fabsf(by), 0x1p-19f, fabsf(ay) * (255.0f * 0x1p-20f));
    const float ez = __builtin_fmaf(fabsf(bz), 0x1p-19f, fabsf(az) * (255.0f * 0x1p-20f));
    // near and far bounds of the four children (qlo words q1.x/y/z, qhi
    // words q1.w, q2.x, q2.y), by the sign of the direction
    const bool px = !(inv.x < 0.0f), py = !(inv.y < 0.0f), pz = !(inv.z < 0.0f);
    const uint32_t nxw = px ? q1.x : q1.w, fxw = px ? q1.w : q1.x;
    const uint32_t nyw = py ? q1.y : q2.x, fyw = py ? q2.x : q1.y;
    const uint32_t nzw = pz ? q1.z : q2.y, fzw = pz ? q2.y : q1.z;
    const float nbx = bx - ex, fbx = bx + ex;
    const float nby = by - ey, fby = by + ey;
    const float nbz = bz - ez, fbz = bz + ez;
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float nx = __builtin_fmaf(byte_of(nxw, i), ax, nbx), fx = __builtin_fmaf(byte_of(fxw, i), ax, fbx);
        const float ny = __builtin_fmaf(byte_of(nyw, i), ay, nby), fy = __builtin_fmaf(byte_of(fyw, i), ay, fby);
        const float nz = __builtin_fmaf(byte_of(nzw, i), az, nbz), fz = __builtin_fmaf(byte_of(fzw, i), az, fbz);
        const float tnear = fmaxf(fmaxf(nx, ny), fmaxf(nz, 0.0f));
        const float tfar = fminf(fminf(fx, fy), fminf(fz, tmax));
        tn[i] = tnear;
        const bool hit = child[i] != kEmpty && tnear <= tfar;
        mask |= hit ? (1u << i) : 0u;
    }
    return mask;
}

RP_DEV void load_tri(const Scene &S, uint32_t p, F3 &a, F3 &b, F3 &c, uint32_t &id) {
    float4 t0 = S.tris[3 * p], t1 = S.tris[3 * p + 1], t2 = S.tris[3 * p + 2];
    a = {t0.x, t0.y, t0.z};
    b = {t1.x, t1.y, t1.z};
    c = {t2.x, t2.y, t2.z};
    id = __float_as_uint(t0.w);
}

constexpr float kCutGrowth = 1.0f + 0x1p-16f;

struct TraversalCounts {
    uint32_t nodes, tris;
    uint32_t spills; // pushes beyond the LDS short stack (into the per-lane scratch overflow)
};

// Compare-exchange on (key, id) pairs; equal keys are never exchanged.
RP_DEV void cswap(float &ka, uint32_t &ia, float &kb, uint32_t &ib) {
    const bool s = kb < ka;
    const float k = s ? kb : ka;
    kb = s ? ka : kb;
    ka = k;
    const uint32_t i = s ? ib : ia;
    ib = s ? ia : ib;
    ia = i;
}

// Closest hit (smallest t, ties to the smallest scene triangle id).  With
// kCount, node fetches, triangle tests and stack pushes that overflow into
// scratch are tallied (instrumented replay).
template <bool kAny, bool kCount = false>
RP_DEV bool traverse(const Scene &S, Stack &st, F3 o, F3 d, float tmax, Hit &h, TraversalCounts *tc = nullptr) {
    h.prim = kEmpty;
    h.id = kEmpty;
    h.t = __builtin_inff();
    if (S.node_count == 0) return false;
    F3 inv = {slab_inv(d.x), slab_inv(d.y), slab_inv(d.z)};
    // Boxes and stack entries are culled beyond `cut` = min(tmax, h.t)
    // grown by 2^-16: a box's computed entry distance and a triangle's
    // computed hit distance round differently, and a triangle that ties the
    // closest hit (a shared edge) must still be reached for the smallest-id
    // rule, whatever the visiting order.  Growing the cut only visits more.
    float cut = tmax * kCutGrowth;
    st.sp = 0;
    uint32_t cur = 0;
    for (;;) {
        if (cur & kLeaf) {
            uint32_t first = cur & ((1u << 27) - 1), count = (cur >> 27) & 15u;
            if (kCount) tc->tris += count;
            for (uint32_t k = 0; k < count; ++k) {
                F3 a, b, c;
                uint32_t id;
                load_tri(S, first + k, a, b, c, id);
                float t, b0, b1, b2;
                if (intersect_triangle(o, d, tmax, a, b, c, t, b0, b1, b2) && (t < h.t || (t == h.t && id < h.id))) {
                    h.prim = first + k;
                    h.id = id;
                    h.t = t;
                    h.b0 = b0;
                    h.b1 = b1;
                    h.b2 = b2;
                    if (kAny) return true;
                    cut = t * kCutGrowth; // t <= tmax
                }
            }
        } else {
            float tn[4];
            uint32_t ch[4];
            if (kCount) tc->nodes += 1;
            uint32_t mask = test_children(S.nodes, cur, o, inv, cut, tn, ch);
            if (mask) {
                // Order hit children by entry distance with a 4-input sorting
                // network on registers (no dynamically indexed arrays, which
                // would live in scratch).  A hit child's entry distance is
                // finite, so the misses (key +inf) sort behind all hits.  Any
                // hit uses the same near-to-far order: occluders near the
                // origin are found first (measured 3 % faster on the 1080p
                // terrain's first shadow bounce than child order).
                const float inf = __builtin_inff();
                float k0 = (mask & 1u) ? tn[0] : inf, k1 = (mask & 2u) ? tn[1] : inf;
                float k2 = (mask & 4u) ? tn[2] : inf, k3 = (mask & 8u) ? tn[3] : inf;
                uint32_t i0 = ch[0], i1 = ch[1], i2 = ch[2], i3 = ch[3];
                cswap(k0, i0, k1, i1);
                cswap(k2, i2, k3, i3);
                cswap(k0, i0, k2, i2);
                cswap(k1, i1, k3, i3);
                cswap(k1, i1, k2, i2);
                // push the farther hits far-to-near through one push site
                for (int j = __builtin_popcount(mask) - 1; j >= 1; --j) {
                    if (kCount && st.sp >= kShortStack) tc->spills += 1;
                    st.push(j == 3 ? i3 : (j == 2 ? i2 : i1), j == 3 ? k3 : (j == 2 ? k2 : k1));
                }
                cur = i0;
                continue;
            }
        }
        // pop the next entry that can still contain a closer hit
        bool found = false;
        while (st.sp > 0) {
            uint32_t n;
            float t;
            st.pop(n, t);
            if (t <= cut) {
                cur = n;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return h.prim != kEmpty;
}

struct SurfacePoint {
    F3 p, n, err;
};
RP_DEV SurfacePoint surface_point(const Scene &S, const Hit &h) {
    F3 a, b, c;
    uint32_t id;
    load_tri(S, h.prim, a, b, c, id);
    SurfacePoint sp;
    sp.p = h.b0 * a + h.b1 * b + h.b2 * c;
    F3 abs_sum = vabs(h.b0 * a) + vabs(h.b1 * b) + vabs(h.b2 * c);
    sp.err = gamma_n(7) * abs_sum;
    sp.n = normalize(cross(a - c, b - c));
    return sp;
}
RP_DEV F3 offset_origin(F3 p, F3 err, F3 n, F3 w) {
    float dd = dot(vabs(n), err);
    F3 off = dd * n;
    if (dot(w, n) < 0) off = -off;
    F3 po = p + off;
    if (off.x > 0) po.x = next_up(po.x); else if (off.x < 0) po.x = next_down(po.x);
    if (off.y > 0) po.y = next_up(po.y); else if (off.y < 0) po.y = next_down(po.y);
    if (off.z > 0) po.z = next_up(po.z); else if (off.z < 0) po.z = next_down(po.z);
    return po;
}
struct Frame {
    F3 x, y, z;
    RP_DEV F3 local(F3 v) const { return {dot(v, x), dot(v, y), dot(v, z)}; }
    RP_DEV F3 world(F3 v) const { return x * v.x + y * v.y + z * v.z; }
};
RP_DEV Frame frame_from_z(F3 n) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    Frame f;
    f.x = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    f.y = {b, sign + n.y * n.y * a, -n.y};
    f.z = n;
    return f;
}

// =================================================================== BSDF
struct Lane {
    uint32_t kind;
    float r, eta, k, ax, ay;
    bool smooth;
};
RP_DEV Lane prepare_lane(const Scene &S, const rp_material &m, float lambda) {
    const uint32_t kind = m.kind;
    Lane l;
    l.kind = kind;
    l.r = 0;
    l.eta = 1;
    l.k = 0;
    l.ax = l.ay = 0;
    l.smooth = false;
    if (kind == RP_MATERIAL_DIFFUSE) {
        l.r = spectrum_eval(m.reflectance, lambda, S.spectra);
        return l;
    }
    l.eta = spectrum_eval(m.eta, lambda, S.spectra);
    if (kind == RP_MATERIAL_CONDUCTOR) l.k = spectrum_eval(m.k, lambda, S.spectra);
    float u = m.roughness_u, v = m.roughness_v;
    l.ax = m.remap_roughness ? __builtin_sqrtf(u) : u;
    l.ay = m.remap_roughness ? __builtin_sqrtf(v) : v;
    if (kind == RP_MATERIAL_DIELECTRIC && l.eta == 0.0f) l.eta = 1.0f;
    l.smooth = omax(l.ax, l.ay) < 1e-3f || (kind == RP_MATERIAL_DIELECTRIC && l.eta == 1.0f);
    if (!l.smooth) {
        l.ax = omax(l.ax, 1e-4f);
        l.ay = omax(l.ay, 1e-4f);
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
RP_DEV Sum2 two_sum(float a, float b) {
    const float s = a + b, bv = s - a;
    return {s, (a - (s - bv)) + (b - bv)};
}
RP_DEV bool finite3(F3 v) { return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z); }
// v / |v| after scaling by the largest component (false: zero or not finite).
RP_DEV bool unit3(F3 v, F3 &out) {
    const float s = omax(omax(fabsf(v.x), fabsf(v.y)), fabsf(v.z));
    if (!(s > 0) || !__builtin_isfinite(s)) return false;
    v = v * (1.0f / s);
    out = v * (1.0f / __builtin_sqrtf(dot(v, v)));
    return finite3(out);
}
// |a|^2 / |b|^2 of two complex numbers, scaled by b's larger component.
RP_DEV bool ratio_norm(float ar, float ai, float br, float bi, float &out) {
    const float s = omax(fabsf(br), fabsf(bi));
    if (!(s > 0) || !__builtin_isfinite(s) || !__builtin_isfinite(ar) || !__builtin_isfinite(ai)) return false;
    const float r = 1.0f / s;
    ar = ar * r, ai = ai * r, br = br * r, bi = bi * r;
    out = (ar * ar + ai * ai) / (br * br + bi * bi);
    return __builtin_isfinite(out);
}
// Conductor Fresnel reflectance at cos(theta) for eta + i k: the complex
// Snell root sqrt(n^2 - sin^2) with the three squares and their sum
// compensated (shaders/conductor.h:57-98).
RP_DEV float fresnel_complex(float cosine, float eta, float k) {
    if (!__builtin_isfinite(cosine) || !__builtin_isfinite(eta) || !__builtin_isfinite(k) || eta < 0 || k < 0 ||
        (eta == 0 && k == 0))
        return __builtin_nanf("");
    const float c = omin(omax(cosine, 0.0f), 1.0f);
    if (eta == 1 && k == 0) return 0.0f;
    if (c == 0) return 1.0f;
    const float n2x = eta * eta - k * k, n2y = 2 * eta * k;
    if (!__builtin_isfinite(n2x) || !__builtin_isfinite(n2y)) return __builtin_nanf("");
    const float eta2 = eta * eta, k2 = k * k, c2 = c * c;
    const Sum2 ab = two_sum(eta2, -k2), cd = two_sum(c2, -1.0f), total = two_sum(ab.s, cd.s);
    const float correction = ab.e + cd.e + total.e + __builtin_fmaf(eta, eta, -eta2) - __builtin_fmaf(k, k, -k2) +
                             __builtin_fmaf(c, c, -c2);
    const float zx = total.s + correction, zy = n2y;
    const float scale = omax(fabsf(zx), fabsf(zy));
    float radius = 0.0f;
    if (scale != 0) {
        const float r = 1.0f / scale, sx = zx * r, sy = zy * r;
        radius = scale * __builtin_sqrtf(sx * sx + sy * sy);
    }
    if (!__builtin_isfinite(radius)) return __builtin_nanf("");
    float qr, qi;
    if (zx >= 0) {
        qr = __builtin_sqrtf(omax(0.0f, 0.5f * radius + 0.5f * zx));
        qi = qr > 0 ? zy / (2 * qr) : 0.0f;
    } else {
        qi = __builtin_sqrtf(omax(0.0f, 0.5f * radius - 0.5f * zx));
        qr = qi > 0 ? zy / (2 * qi) : 0.0f;
    }
    float rs, rp;
    if (!ratio_norm(c - qr, 0.0f - qi, c + qr, qi, rs) ||
        !ratio_norm(n2x * c - qr, n2y * c - qi, n2x * c + qr, n2y * c + qi, rp))
        return __builtin_nanf("");
    const float f = (rs + rp) * 0.5f;
    return __builtin_isfinite(f) && f >= 0 ? f : __builtin_nanf("");
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
RP_DEV Optics optics(float cosine, float eta) {
    Optics o = {0.0f, 1.0f, 0.0f, false, false};
    if (!__builtin_isfinite(cosine)) return o;
    o.etap = cosine < 0 ? 1 / eta : eta;
    const float c = omin(fabsf(cosine), 1.0f);
    if (!(o.etap > 0) || !__builtin_isfinite(o.etap)) return o;
    if (o.etap == 1) {
        o.ct = c;
        o.ok = true;
        return o;
    }
    const float c2 = c * c, e2 = eta * eta;
    float disc;
    if (cosine < 0) {
        const Sum2 a = two_sum(1.0f, -e2);
        disc = __builtin_fmaf(e2, c2, a.s) +
               (a.e + __builtin_fmaf(eta, eta, -e2) * (c2 - 1) + e2 * __builtin_fmaf(c, c, -c2));
    } else {
        const Sum2 a = two_sum(e2, -1.0f), b = two_sum(a.s, c2);
        disc = b.s + (a.e + b.e + __builtin_fmaf(eta, eta, -e2) + __builtin_fmaf(c, c, -c2));
    }
    if (!__builtin_isfinite(disc)) return o;
    o.tir = disc <= 0;
    if (o.tir) {
        o.F = 1.0f;
        o.ok = true;
        return o;
    }
    o.ct = __builtin_sqrtf(disc) / (cosine < 0 ? 1.0f : eta);
    const float rp = (o.etap * c - o.ct) / (o.etap * c + o.ct);
    const float rs = (c - o.etap * o.ct) / (c + o.etap * o.ct);
    o.F = 0.5f * (rp * rp + rs * rs);
    o.ok = __builtin_isfinite(o.ct) && __builtin_isfinite(o.F) && o.F >= 0 && o.F <= 1;
    return o;
}
RP_DEV float fresnel_dielectric(float cosine, float eta) {
    const Optics o = optics(cosine, eta);
    return o.ok ? o.F : __builtin_nanf("");
}
// c / etap - ct written with the authored index on both sides
// (shaders/dielectric.h:125-130).
RP_DEV float refraction_coefficient(float eta, float co, float ct) {
    const float c = fabsf(co);
    return co < 0 ? ((eta - 1) * (eta + 1)) / (eta * c + ct) : (((1 - eta) * (1 + eta)) / eta) / (c + eta * ct);
}
// Trowbridge-Reitz D with PBRT's grazing cutoff, and Smith Lambda.
RP_DEV float ggx_d(const Lane &l, F3 m) {
    if (m.z * m.z * m.z * m.z < 1e-16f) return 0.0f;
    F3 v = {m.x / l.ax, m.y / l.ay, m.z};
    float t = dot(v, v);
    const float den = kPi * l.ax * l.ay * t * t;
    if (!__builtin_isfinite(den) || !(den > 0)) return __builtin_nanf("");
    const float d = 1.0f / den;
    return d > 0 ? d : __builtin_nanf("");
}
RP_DEV float ggx_lambda(const Lane &l, F3 w) {
    const F3 v = {1.0f, l.ax * w.x / w.z, l.ay * w.y / w.z};
    const float s = omax(omax(fabsf(v.x), fabsf(v.y)), fabsf(v.z));
    return 0.5f * (s * length(v * (1.0f / s)) - 1.0f);
}
// Heitz 2018 visible-normal sampling; the height over the warped disk keeps
// 1 - u at the disk's edge (shaders/conductor.h:153-178).  False: invalid.
RP_DEV bool ggx_sample_wm(const Lane &l, F3 w, float u0, float u1, F3 &m) {
    F3 wh;
    if (!unit3(F3{l.ax * w.x, l.ay * w.y, w.z}, wh)) return false;
    if (wh.z < 0) wh = -wh;
    F3 t1 = {1.0f, 0.0f, 0.0f};
    if (wh.z < 0.99999f && !unit3(F3{-wh.y, wh.x, 0.0f}, t1)) return false;
    const F3 t2 = cross3(wh, t1);
    const float r = __builtin_sqrtf(u0);
    float sn, cs;
    sincos_turns(u1, sn, cs);
    const float dx = r * cs, old_y = r * sn;
    const float h = __builtin_sqrtf(omax(0.0f, 1.0f - dx * dx)), a = (1.0f + wh.z) * 0.5f;
    const float dy = (1.0f - a) * h + a * old_y;
    const float old_minus_h = old_y > 0 ? -(1.0f - u0) / (h + old_y) : old_y - h;
    const float z = __builtin_sqrtf(omax(0.0f, (1.0f - u0) + (1.0f - a) * old_minus_h * (old_y + dy)));
    const F3 nh = dx * t1 + dy * t2 + z * wh;
    return unit3(F3{l.ax * nh.x, l.ay * nh.y, omax(1e-6f, nh.z)}, m);
}
struct Eval {
    float f, pdf;
};
// Rough dielectric at (wo, wi) with its half vector (shaders/dielectric.h:72-124).
RP_DEV Eval dielectric_eval(const Lane &l, F3 wo, F3 wi) {
    const float nan = __builtin_nanf("");
    const bool reflection = (wo.z > 0) == (wi.z > 0), entering = wo.z > 0;
    const F3 h = reflection ? wo + wi
                            : (entering ? F3{__builtin_fmaf(l.eta, wi.x, wo.x), __builtin_fmaf(l.eta, wi.y, wo.y),
                                             __builtin_fmaf(l.eta, wi.z, wo.z)}
                                        : F3{__builtin_fmaf(l.eta, wo.x, wi.x), __builtin_fmaf(l.eta, wo.y, wi.y),
                                             __builtin_fmaf(l.eta, wo.z, wi.z)});
    if (h.x == 0 && h.y == 0 && h.z == 0) return {0.0f, 0.0f};
    F3 m;
    if (!unit3(h, m)) return {nan, nan};
    if (m.z < 0) m = -m;
    const float co = dot(wo, m), ci = dot(wi, m);
    if (co * wo.z < 0 || ci * wi.z < 0) return {0.0f, 0.0f};
    const Optics o = optics(co, l.eta);
    if (!o.ok) return {nan, nan};
    const float lo = ggx_lambda(l, wo), li = ggx_lambda(l, wi), d = ggx_d(l, m);
    if (!__builtin_isfinite(lo) || !__builtin_isfinite(li) || !__builtin_isfinite(d)) return {nan, nan};
    if (d == 0) return {0.0f, 0.0f};
    if (reflection)
        return {d * o.F / (4 * fabsf(wo.z) * fabsf(wi.z) * (1 + lo + li)), d * o.F / (4 * fabsf(wo.z) * (1 + lo))};
    const float den = dot(h, m) / (entering ? l.eta : 1.0f);
    const float jacobian = fabsf(ci) / (den * den);
    const float pdf = d * fabsf(co) / (fabsf(wo.z) * (1 + lo)) * jacobian * (1 - o.F);
    const float scale = entering ? 1 / (l.eta * l.eta) : l.eta * l.eta;
    return {d * (1 - o.F) * fabsf(ci * co) * scale / (fabsf(wo.z) * fabsf(wi.z) * den * den * (1 + lo + li)), pdf};
}
RP_DEV Eval eval_lane(const Lane &l, F3 wo, F3 wi) {
    Eval e = {0.0f, 0.0f};
    if (l.kind == RP_MATERIAL_DIFFUSE) {
        if (!(wo.z * wi.z > 0)) return e;
        e.f = l.r * kInvPi;
        e.pdf = fabsf(wi.z) * kInvPi;
        return e;
    }
    if (l.smooth || wo.z == 0.0f || wi.z == 0.0f) return e;
    if (l.kind == RP_MATERIAL_CONDUCTOR) {
        if ((wo.z > 0) != (wi.z > 0)) return e;
        F3 m;
        if (!unit3(wo + wi, m)) return {__builtin_nanf(""), __builtin_nanf("")};
        const float F = fresnel_complex(fabsf(dot(wo, m)), l.eta, l.k);
        const float lo = ggx_lambda(l, wo), li = ggx_lambda(l, wi), D = ggx_d(l, m);
        e.f = D * F / (4.0f * fabsf(wo.z) * fabsf(wi.z) * (1.0f + lo + li));
        e.pdf = D / (4.0f * fabsf(wo.z) * (1.0f + lo));
        return e;
    }
    return dielectric_eval(l, wo, wi);
}
// eval_lane for the four hero lanes of one material: the terms that depend
// only on the directions and the roughness (half vector, Smith Lambda, D,
// the denominators) once, the wavelength-dependent factors (Fresnel, albedo,
// a dielectric's refracted half vector) per lane, and a lane whose optical
// constants equal lane 0's takes lane 0's result.  Each e[k] is bit for bit
// eval_lane(l[k], wo, wi): the same operations on the same values.
RP_DEV void eval_hero(const Lane l[4], F3 wo, F3 wi, Eval e[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) e[k] = {0.0f, 0.0f};
    const uint32_t kind = l[0].kind;
    if (kind == RP_MATERIAL_DIFFUSE) {
        if (!(wo.z * wi.z > 0)) return;
        const float pdf = fabsf(wi.z) * kInvPi;
#pragma unroll
        for (int k = 0; k < 4; ++k) e[k] = {l[k].r * kInvPi, pdf};
        return;
    }
    if (wo.z == 0.0f || wi.z == 0.0f) return;
    if (kind == RP_MATERIAL_CONDUCTOR) {
        // a conductor lane is smooth by its roughness alone: all lanes alike
        if (l[0].smooth || (wo.z > 0) != (wi.z > 0)) return;
        F3 m;
        if (!unit3(wo + wi, m)) {
            const float nan = __builtin_nanf("");
#pragma unroll
            for (int k = 0; k < 4; ++k) e[k] = {nan, nan};
            return;
        }
        const float c = fabsf(dot(wo, m));
        const float lo = ggx_lambda(l[0], wo), li = ggx_lambda(l[0], wi), D = ggx_d(l[0], m);
        const float den = 4.0f * fabsf(wo.z) * fabsf(wi.z) * (1.0f + lo + li);
        const float pdf = D / (4.0f * fabsf(wo.z) * (1.0f + lo));
        const float F0 = fresnel_complex(c, l[0].eta, l[0].k);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float F = (k == 0 || (l[k].eta == l[0].eta && l[k].k == l[0].k)) ? F0
                                                                                    : fresnel_complex(c, l[k].eta, l[k].k);
            e[k] = {D * F / den, pdf};
        }
        return;
    }
    if (l[0].smooth) { // by roughness all lanes are; by an index of 1 only that lane
#pragma unroll
        for (int k = 1; k < 4; ++k)
            if (!l[k].smooth) e[k] = dielectric_eval(l[k], wo, wi);
        return;
    }
    const float nan = __builtin_nanf("");
    if ((wo.z > 0) == (wi.z > 0)) {
        // dielectric reflection: the half vector, Lambda and D are every
        // lane's own; only the Fresnel term differs
        const F3 h = wo + wi;
        if (h.x == 0 && h.y == 0 && h.z == 0) return;
        F3 m;
        if (!unit3(h, m)) {
#pragma unroll
            for (int k = 0; k < 4; ++k) e[k] = {nan, nan};
            return;
        }
        if (m.z < 0) m = -m;
        const float co = dot(wo, m), ci = dot(wi, m);
        if (co * wo.z < 0 || ci * wi.z < 0) return;
        const float lo = ggx_lambda(l[0], wo), li = ggx_lambda(l[0], wi), d = ggx_d(l[0], m);
        const bool finite = __builtin_isfinite(lo) && __builtin_isfinite(li) && __builtin_isfinite(d);
        const float den_f = 4 * fabsf(wo.z) * fabsf(wi.z) * (1 + lo + li), den_p = 4 * fabsf(wo.z) * (1 + lo);
        const Optics o0 = optics(co, l[0].eta);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (l[k].smooth) continue;
            const Optics o = (k == 0 || l[k].eta == l[0].eta) ? o0 : optics(co, l[k].eta);
            if (!o.ok || !finite) e[k] = {nan, nan};
            else if (d != 0) e[k] = {d * o.F / den_f, d * o.F / den_p};
        }
        return;
    }
    e[0] = dielectric_eval(l[0], wo, wi);
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (!l[k].smooth) e[k] = l[k].eta == l[0].eta ? e[0] : dielectric_eval(l[k], wo, wi);
}

struct Sampled {
    F3 wi;
    float pdf, eta;
    bool delta, transmission, ok;
};
// Refraction of wo about the microfacet normal m with the rationalised
// coefficient (shaders/dielectric.h:176-186); etap is the relative index of
// the side entered.  False on total internal reflection or an invalid form.
RP_DEV bool refract(F3 wo, F3 m, float eta, float &etap, F3 &wt) {
    const float co = dot(wo, m);
    const Optics o = optics(co, eta);
    if (!o.ok || o.tir) return false;
    const float coefficient = eta == 1 ? 0.0f : refraction_coefficient(eta, co, o.ct);
    wt = (co < 0 ? (-eta) * wo : (-wo) / eta) + coefficient * (co < 0 ? -m : m);
    etap = o.etap;
    return finite3(wt);
}
RP_DEV F3 cosine_hemisphere(float u0, float u1) {
    float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (ox != 0.0f || oy != 0.0f) {
        float r, turns;
        if (fabsf(ox) > fabsf(oy)) {
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
    float z = __builtin_sqrtf(omax(0.0f, 1.0f - dx * dx - dy * dy));
    return {dx, dy, z};
}
RP_DEV Sampled sample_lane(const Lane &l, F3 wo, float uc, float u0, float u1) {
    Sampled s;
    s.wi = F3{0.0f, 0.0f, 0.0f};
    s.pdf = 0.0f;
    s.eta = 1.0f;
    s.delta = s.transmission = s.ok = false;
    if (wo.z == 0.0f) return s;
    if (l.kind == RP_MATERIAL_DIFFUSE) {
        F3 wi = cosine_hemisphere(u0, u1);
        if (wo.z < 0) wi.z = -wi.z;
        s.wi = wi;
        s.pdf = fabsf(wi.z) * kInvPi;
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
        F3 m;
        if (!ggx_sample_wm(l, wo, u0, u1, m)) return s;
        const F3 wi = -wo + 2.0f * dot(wo, m) * m;
        if (!finite3(wi) || wi.z == 0.0f || (wo.z > 0) != (wi.z > 0)) return s;
        const float lo = ggx_lambda(l, wo), D = ggx_d(l, m);
        if (!(D > 0)) return s;
        s.wi = wi;
        s.pdf = D / (4.0f * fabsf(wo.z) * (1.0f + lo));
        s.ok = __builtin_isfinite(s.pdf) && s.pdf > 0;
        return s;
    }
    // dielectric (shaders/dielectric.h:131-228): the lobe from the Fresnel
    // reflectance at the sampled microfacet normal (the geometric normal when
    // smooth)
    F3 m = {0.0f, 0.0f, 1.0f};
    if (!l.smooth && !ggx_sample_wm(l, wo, u0, u1, m)) return s;
    const float co = dot(wo, m);
    const Optics o = optics(co, l.eta);
    if (!o.ok) return s;
    const bool reflection = uc < o.F;
    F3 wi;
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
        if (!__builtin_isfinite(lo) || !__builtin_isfinite(li) || !__builtin_isfinite(D) || D == 0) return s;
        if (reflection) pdf = D * o.F / (4 * fabsf(wo.z) * (1 + lo));
        else {
            const float den = refraction_coefficient(l.eta, co, o.ct);
            pdf = D * fabsf(co) * o.ct * (1 - o.F) / (fabsf(wo.z) * (1 + lo) * den * den);
        }
    }
    s.wi = wi;
    s.pdf = pdf;
    s.delta = l.smooth;
    s.transmission = !reflection;
    s.eta = reflection ? 1.0f : o.etap;
    s.ok = __builtin_isfinite(pdf) && pdf > 0;
    return s;
}
// A companion lane's mass and density for the hero's Dirac direction wi: the
// lane keeps mass only if its own specular direction is bit for bit wi.
RP_DEV void dirac_lane(const Lane &l, F3 wo, F3 wi, float u0, float u1, float &mass, float &pdf) {
    mass = pdf = 0.0f;
    if (wo.z == 0.0f || wi.z == 0.0f) return;
    bool transmission = wo.z * wi.z < 0;
    if (l.kind == RP_MATERIAL_CONDUCTOR && l.smooth) {
        if (transmission || !(wi.x == -wo.x && wi.y == -wo.y && wi.z == wo.z)) return;
        pdf = 1.0f;
        mass = fresnel_complex(fabsf(wo.z), l.eta, l.k);
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

// ================================================================= media
struct MediumCoeff {
    float st[4], ss[4];
    float g;
};
RP_DEV MediumCoeff medium_coeff(const Scene &S, uint32_t m, const Wavelengths &wl) {
    MediumCoeff c;
    const rp_medium &md = S.media[m];
    for (int k = 0; k < 4; ++k) {
        float a = spectrum_eval(md.sigma_a, wl.lambda[k], S.spectra);
        float sc = spectrum_eval(md.sigma_s, wl.lambda[k], S.spectra);
        c.ss[k] = sc;
        c.st[k] = a + sc;
    }
    c.g = omin(omax(md.g, -0.99f), 0.99f);
    return c;
}
RP_DEV float hg(float cos_t, float g) {
    float denom = 1.0f + g * g + 2.0f * g * cos_t;
    return kInv4Pi * (1.0f - g * g) / (denom * __builtin_sqrtf(denom));
}
RP_DEV F3 hg_sample(F3 wo, float g, float u0, float u1, float &pdf) {
    float cos_t;
    if (fabsf(g) < 1e-3f) cos_t = 1.0f - 2.0f * u0;
    else {
        float sq = (1.0f - g * g) / (1.0f + g - 2.0f * g * u0);
        cos_t = -1.0f / (2.0f * g) * (1.0f + g * g - sq * sq);
    }
    cos_t = omin(omax(cos_t, -1.0f), 1.0f);
    float sin_t = __builtin_sqrtf(omax(0.0f, 1.0f - cos_t * cos_t));
    float sn, cs;
    sincos_turns(u1, sn, cs);
    Frame f = frame_from_z(wo);
    F3 wi = f.world(F3{sin_t * cs, sin_t * sn, cos_t});
    pdf = hg(cos_t, g);
    return wi;
}
RP_DEV uint32_t outgoing_medium(const rp_object &o, uint32_t current, F3 n, F3 w) {
    if (o.medium_inside == o.medium_outside) return current;
    return dot(n, w) > 0 ? o.medium_outside : o.medium_inside;
}

// ================================================================ lights
struct LightSample {
    F3 p, n, err, wi;
    float le[4];
    float pdf, dist;
    bool delta, infinite, ok;
};
// --- light selection (shaders/light_sampling.h:19-127; PBRT BVHLightSampler,
// PowerLightSampler, UniformLightSampler).  A receiving point p with normal
// n (zero in media).
RP_DEV F3 node_vec(const float v[3]) { return F3{v[0], v[1], v[2]}; }
RP_DEV bool nonzero(F3 n) { return n.x != 0.0f || n.y != 0.0f || n.z != 0.0f; }
RP_DEV float light_importance(const LightNode &b, F3 p, F3 n) {
    const F3 delta = p - node_vec(b.center);
    const float r2 = b.r2, d2 = dot(delta, delta);
    const float radius = b.radius, dist = __builtin_sqrtf(d2);
    const float cb = d2 <= r2 ? -1.0f : __builtin_sqrtf(omax(0.0f, 1.0f - r2 / d2));
    const float sb = __builtin_sqrtf(omax(0.0f, 1.0f - cb * cb));
    const F3 wi = dist > 0 ? delta * (1.0f / dist) : F3{0, 0, 1};
    float cw = oclamp(dot(node_vec(b.dir), wi), -1.0f, 1.0f);
    if (b.flags & 1u) cw = fabsf(cw);
    const float sw = __builtin_sqrtf(omax(0.0f, 1.0f - cw * cw));
    const float co = b.cos_outer, so = b.sin_outer;
    const float cx = cw > co ? 1.0f : cw * co + sw * so, sx = cw > co ? 0.0f : sw * co - cw * so;
    const float cp = cx > cb ? 1.0f : cx * cb + sx * sb;
    if (cp <= b.cos_emission) return 0.0f;
    float value = b.power * cp / omax(d2, omax(radius, 1.17549435e-38f));
    if (nonzero(n)) {
        const float ci = oclamp(fabsf(dot(wi, n)), 0.0f, 1.0f);
        const float si = __builtin_sqrtf(omax(0.0f, 1.0f - ci * ci));
        value = value * (ci > cb ? 1.0f : ci * cb + si * sb);
    }
    return omax(0.0f, value);
}
struct Selected {
    uint32_t light;
    float pmf;
};
RP_DEV Selected light_select(const Scene &S, F3 p, F3 n, float u) {
    const uint32_t count = S.light_count;
    if (count == 0) return {kEmpty, 0.0f};
    u = oclamp(u, 0.0f, 0x1.fffffep-1f);
    if (S.light_sampler == RP_LIGHT_SAMPLER_UNIFORM) {
        const uint32_t i = uint32_t(u * float(count));
        return {i < count - 1 ? i : count - 1, 1.0f / float(count)};
    }
    if (S.light_sampler == RP_LIGHT_SAMPLER_POWER) {
        const float v = u * float(count);
        uint32_t i = uint32_t(v);
        i = i < count - 1 ? i : count - 1;
        const uint32_t id = v - float(i) < S.lights[i].threshold ? i : S.lights[i].alias;
        return {id, S.lights[id].pmf};
    }
    const uint32_t inf = S.env_light != kEmpty ? 1u : 0u, tree = S.light_node_count ? 1u : 0u;
    const float infinite = float(inf) / float(inf + tree);
    if (u < infinite) return {S.env_light, 1.0f / float(inf + tree)};
    if (!tree) return {kEmpty, 0.0f};
    u = omin((u - infinite) / (1.0f - infinite), 0x1.fffffep-1f);
    float pmf = 1.0f - infinite;
    uint32_t i = 0;
    // The chosen child's record and importance are carried down, so a level
    // costs one round trip (the two children's records, loaded together).
    // A chosen child has positive importance; only a root leaf is tested.
    LightNode node = S.light_nodes[0];
    if (node.light != kEmpty)
        return light_importance(node, p, n) <= 0.0f ? Selected{kEmpty, 0.0f} : Selected{node.light, pmf};
    for (uint32_t steps = 0; steps < S.light_node_count; ++steps) {
        const LightNode na = S.light_nodes[i + 1], nb = S.light_nodes[node.child];
        const float a = light_importance(na, p, n);
        const float b = light_importance(nb, p, n);
        const float sum = a + b;
        if (!(sum > 0.0f)) return {kEmpty, 0.0f};
        const float q = a / sum;
        if (u < q) {
            pmf = pmf * q;
            u = omin(u / q, 0x1.fffffep-1f);
            i = i + 1;
            node = na;
        } else {
            pmf = pmf * (1.0f - q);
            u = omin((u - q) / (1.0f - q), 0x1.fffffep-1f);
            i = node.child;
            node = nb;
        }
        if (node.light != kEmpty) return {node.light, pmf};
    }
    return {kEmpty, 0.0f};
}
RP_DEV float light_pmf(const Scene &S, F3 p, F3 n, uint32_t id) {
    const uint32_t count = S.light_count;
    if (id >= count) return 0.0f;
    if (S.light_sampler == RP_LIGHT_SAMPLER_UNIFORM) return 1.0f / float(count);
    if (S.light_sampler == RP_LIGHT_SAMPLER_POWER) return S.lights[id].pmf;
    const uint32_t inf = S.env_light != kEmpty ? 1u : 0u, tree = S.light_node_count ? 1u : 0u;
    if (id == S.env_light) return 1.0f / float(inf + tree);
    uint32_t i = S.lights[id].leaf;
    if (i >= S.light_node_count) return 0.0f;
    if (light_importance(S.light_nodes[i], p, n) <= 0.0f) return 0.0f;
    float pmf = 1.0f - float(inf) / float(inf + tree);
    for (uint32_t steps = 0; steps < S.light_node_count; ++steps) {
        const uint32_t parent = S.light_nodes[i].parent;
        if (parent == kEmpty) return pmf;
        const float a = light_importance(S.light_nodes[parent + 1], p, n);
        const float b = light_importance(S.light_nodes[S.light_nodes[parent].child], p, n);
        const float sum = a + b;
        if (!(sum > 0.0f)) return 0.0f;
        pmf = pmf * (i == parent + 1 ? a / sum : b / sum);
        i = parent;
    }
    return 0.0f;
}

// --- quad emitters: the spherical-rectangle proposal (Urena, Fajardo, King
// 2013; src/light_geometry_sampling_shared.h:174-260,
// shaders/light_geometry_sampling.h:118-209), warped by the receiver's corner
// cosines (bilinear) when a normal is given; uniform over the area when the
// rectangle subtends 1e-4 sr or less.
RP_DEV F3 unit(F3 v) { return v * (1.0f / __builtin_sqrtf(dot(v, v))); }
RP_DEV float angle_between(F3 a, F3 b) {
    return dot(a, b) < 0 ? kPi - 2.0f * asin_f(omin(1.0f, length(a + b) / 2.0f))
                         : 2.0f * asin_f(omin(1.0f, length(a - b) / 2.0f));
}
RP_DEV float spherical_triangle_area(F3 a, F3 b, F3 c) {
    return fabsf(2.0f * atan2_f(dot(a, cross3(b, c)), 1.0f + dot(a, b) + dot(a, c) + dot(b, c)));
}
struct Bilinear {
    float a, b, c, d; // weights at (0,0), (1,0), (0,1), (1,1)
};
RP_DEV Bilinear bilinear_normalized(Bilinear w) {
    const float m = omax(omax(w.a, w.b), omax(w.c, w.d));
    return m > 0 ? Bilinear{w.a / m, w.b / m, w.c / m, w.d / m} : Bilinear{1, 1, 1, 1};
}
RP_DEV float linear_warp(float u, float a, float b) {
    if (u == 0.0f && a == 0.0f) return 0.0f;
    return omin(0x1.fffffep-1f, u * (a + b) / (a + __builtin_sqrtf((1.0f - u) * a * a + u * b * b)));
}
RP_DEV void bilinear_warp(float &u0, float &u1, Bilinear w) {
    w = bilinear_normalized(w);
    const float y = linear_warp(u1, w.a + w.b, w.c + w.d);
    const float x = linear_warp(u0, (1.0f - y) * w.a + y * w.c, (1.0f - y) * w.b + y * w.d);
    u0 = x;
    u1 = y;
}
RP_DEV float bilinear_density(float x, float y, Bilinear w) {
    if (x < 0 || x > 1 || y < 0 || y > 1) return 0.0f;
    w = bilinear_normalized(w);
    return 4.0f * ((1.0f - x) * (1.0f - y) * w.a + x * (1.0f - y) * w.b + (1.0f - x) * y * w.c + x * y * w.d) /
           (w.a + w.b + w.c + w.d);
}
// The rectangle seen from r in its own frame: extents, spherical-rectangle
// constants, solid angle, unit directions to the four corners.
struct RectFrame {
    float x0, x1, y0, y1, z0, b0, b1, g2, g3, area, zs;
    F3 w00, w10, w01, w11;
};
RP_DEV RectFrame rect_frame(const LightEntry &L, F3 r) {
    RectFrame f;
    const F3 delta = node_vec(L.p00) - r;
    f.x0 = dot(delta, node_vec(L.ux));
    f.x1 = f.x0 + L.lx;
    f.y0 = dot(delta, node_vec(L.uy));
    f.y1 = f.y0 + L.ly;
    f.z0 = dot(delta, node_vec(L.uz));
    f.zs = 1.0f;
    if (f.z0 > 0) {
        f.z0 = -f.z0;
        f.zs = -1.0f;
    }
    const F3 v00 = {f.x0, f.y0, f.z0}, v10 = {f.x1, f.y0, f.z0}, v11 = {f.x1, f.y1, f.z0}, v01 = {f.x0, f.y1, f.z0};
    const F3 n0 = unit(cross3(v00, v10)), n2 = unit(cross3(v11, v01)), n3 = unit(cross3(v01, v00));
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
RP_DEV Bilinear rect_cosines(const RectFrame &f, const LightEntry &L, F3 n) {
    const F3 nl = {dot(n, node_vec(L.ux)), dot(n, node_vec(L.uy)), f.zs * dot(n, node_vec(L.uz))};
    return {omax(0.01f, fabsf(dot(f.w00, nl))), omax(0.01f, fabsf(dot(f.w10, nl))), omax(0.01f, fabsf(dot(f.w01, nl))),
            omax(0.01f, fabsf(dot(f.w11, nl)))};
}
RP_DEV void rect_sample(const RectFrame &f, float s0, float s1, float &u, float &v) {
    const float au = s0 * f.area - f.g2 - f.g3;
    const float fu = (cos_f(au) * f.b0 - f.b1) / sin_f(au);
    const float cu = omin(0x1.fffffep-1f,
                          omax(-0x1.fffffep-1f, copysignf(1.0f / __builtin_sqrtf(fu * fu + f.b0 * f.b0), fu)));
    const float xu = omin(f.x1, omax(f.x0, -cu * f.z0 / __builtin_sqrtf(omax(0.0f, 1.0f - cu * cu))));
    const float dd = __builtin_sqrtf(xu * xu + f.z0 * f.z0);
    const float h0 = f.y0 / __builtin_sqrtf(dd * dd + f.y0 * f.y0), h1 = f.y1 / __builtin_sqrtf(dd * dd + f.y1 * f.y1);
    const float hv = h0 + s1 * (h1 - h0);
    const float yv = omin(f.y1, omax(f.y0, hv * dd / __builtin_sqrtf(omax(0.0f, 1.0f - hv * hv))));
    u = oclamp((xu - f.x0) / (f.x1 - f.x0), 0.0f, 1.0f);
    v = oclamp((yv - f.y0) / (f.y1 - f.y0), 0.0f, 1.0f);
}
RP_DEV void rect_inverse(const RectFrame &f, float u, float v, float &s0, float &s1) {
    float xu = f.x0 + u * (f.x1 - f.x0);
    const float yv = f.y0 + v * (f.y1 - f.y0);
    if (xu == 0.0f) xu = 1e-10f;
    const float fusq = omax(0.0f, 1.0f + f.z0 * f.z0 / (xu * xu) - f.b0 * f.b0);
    const float fu = copysignf(__builtin_sqrtf(fusq), xu);
    const float root = __builtin_sqrtf(omax(0.0f, f.b0 * f.b0 - f.b1 * f.b1 + fusq));
    float au = atan2_f(-f.b1 * fu - copysignf(f.b0 * root, fu * f.b0), f.b0 * f.b1 - root * fabsf(fu));
    if (au > 0) au = au - 2.0f * kPi;
    s0 = oclamp((au + f.g2 + f.g3) / f.area, 0.0f, 1.0f);
    const float dd2 = xu * xu + f.z0 * f.z0;
    const float h0 = f.y0 / __builtin_sqrtf(dd2 + f.y0 * f.y0), h1 = f.y1 / __builtin_sqrtf(dd2 + f.y1 * f.y1);
    const float hv = yv / __builtin_sqrtf(dd2 + yv * yv);
    s1 = oclamp((hv - h0) / (h1 - h0), 0.0f, 1.0f);
}
// Area density of quad point q seen from r, as solid-angle density.
RP_DEV float quad_area_pdf(const LightEntry &L, F3 r, F3 q) {
    const F3 delta = q - r;
    const float d2 = dot(delta, delta);
    if (!(d2 > 0)) return 0.0f;
    const float c = fabsf(dot(node_vec(L.uz), delta * (1.0f / __builtin_sqrtf(d2))));
    return c > 0 ? L.inv_area * d2 / c : 0.0f;
}
// Solid-angle density of the quad proposal at rectangle coordinates (u, v),
// point q, from r with receiver normal n.
RP_DEV float quad_pdf(const LightEntry &L, F3 r, F3 n, float u, float v, F3 q) {
    if (L.rect) {
        const RectFrame f = rect_frame(L, r);
        if (f.area > 1e-4f) {
            float s0, s1;
            rect_inverse(f, u, v, s0, s1);
            return (nonzero(n) ? bilinear_density(s0, s1, rect_cosines(f, L, n)) : 1.0f) / f.area;
        }
    }
    return quad_area_pdf(L, r, q);
}
RP_DEV void quad_coords(const LightEntry &L, F3 q, float &u, float &v) {
    const F3 d = q - node_vec(L.p00), ex = node_vec(L.ex), ey = node_vec(L.ey);
    u = oclamp(dot(d, ex) / dot(ex, ex), 0.0f, 1.0f);
    v = oclamp(dot(d, ey) / dot(ey, ey), 0.0f, 1.0f);
}

// --- triangle emitters: the spherical-triangle proposal (Arvo 1995, as
// PBRT v4's SampleSphericalTriangle with the reference's rationalised
// 1 - cos(beta'); src/light_geometry_sampling_shared.h:109-176,
// shaders/light_geometry_sampling.h:114-209) for triangles subtending 3e-4
// to 6.22 sr, warped by the receiver's vertex cosines when a normal is
// given; otherwise uniform over the area, (b1, b2) = (1 - sqrt u0, u1 sqrt u0).
RP_DEV bool positive(F3 v) { return dot(v, v) > 0; }
RP_DEV float tri_solid_angle(F3 p0, F3 p1, F3 p2, F3 r) {
    const F3 a = p0 - r, b = p1 - r, c = p2 - r;
    if (!positive(a) || !positive(b) || !positive(c)) return 0.0f;
    return spherical_triangle_area(unit(a), unit(b), unit(c));
}
RP_DEV bool spherical_tri(float area) { return area >= 3e-4f && area <= 6.22f; }
// Bilinear weights of the vertex cosines: (b, b, a, c) over the square.
RP_DEV Bilinear tri_cosines(F3 p0, F3 p1, F3 p2, F3 r, F3 n) {
    const float wa = omax(0.01f, fabsf(dot(n, unit(p0 - r))));
    const float wb = omax(0.01f, fabsf(dot(n, unit(p1 - r))));
    const float wc = omax(0.01f, fabsf(dot(n, unit(p2 - r))));
    return {wb, wb, wa, wc};
}
// Square sample -> barycentrics (b1, b2) of p1 and p2 on the triangle seen
// from r.  False: a degenerate configuration.
RP_DEV bool tri_sample(F3 p0, F3 p1, F3 p2, F3 r, float u0, float u1, float &b1, float &b2) {
    const F3 a = unit(p0 - r), b = unit(p1 - r), c = unit(p2 - r);
    const F3 nab = cross3(a, b), nca = cross3(c, a);
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
    const float omc = omin(2.0f, omax(0.0f, (-k2 * one_minus_ct - k1 * st) / divisor));
    const float cosbp = 1 - omc;
    const F3 arc = c - dot(c, a) * a;
    if (!positive(arc)) return false;
    const F3 cprime = cosbp * a + __builtin_sqrtf(omax(0.0f, omc * (2 - omc))) * unit(arc);
    const float omct = u1 * 0.5f * dot(cprime - b, cprime - b);
    const float ct = 1 - omct;
    const F3 tangent = cprime - dot(cprime, b) * b;
    if (!positive(tangent)) return false;
    const F3 dir = ct * b + __builtin_sqrtf(omax(0.0f, omct * (2 - omct))) * unit(tangent);
    const F3 e1 = p1 - p0, e2 = p2 - p0, s1 = cross3(dir, e2);
    const float denom = dot(s1, e1);
    if (denom == 0) return false;
    const F3 s = r - p0;
    b1 = omin(1.0f, omax(0.0f, dot(s, s1) / denom));
    b2 = omin(1.0f, omax(0.0f, dot(dir, cross3(s, e1)) / denom));
    if (b1 + b2 > 1) {
        const float sum = b1 + b2;
        b1 = b1 / sum;
        b2 = b2 / sum;
    }
    return __builtin_isfinite(b1) && __builtin_isfinite(b2);
}
// Direction dir from r (a point of the triangle) -> square sample.
RP_DEV void tri_inverse(F3 p0, F3 p1, F3 p2, F3 r, F3 dir, float &s0, float &s1) {
    const F3 a = unit(p0 - r), b = unit(p1 - r), c = unit(p2 - r);
    F3 cp = unit(cross3(cross3(b, dir), cross3(c, a)));
    if (dot(cp, a + c) < 0) cp = -cp;
    s0 = omin(1.0f, omax(0.0f, spherical_triangle_area(a, b, cp) / spherical_triangle_area(a, b, c)));
    s1 = omin(1.0f, omax(0.0f, dot(dir - b, dir - b) / dot(cp - b, cp - b)));
}
// Solid-angle density of the triangle proposal at point q (geometric normal
// ng) from r with receiver normal n.
RP_DEV float tri_pdf(const LightEntry &L, F3 p0, F3 p1, F3 p2, F3 r, F3 n, F3 q, F3 ng) {
    const F3 dv = q - r;
    const float d2 = dot(dv, dv);
    if (!(d2 > 0)) return 0.0f;
    const float area = tri_solid_angle(p0, p1, p2, r);
    if (spherical_tri(area)) {
        float s0, s1;
        tri_inverse(p0, p1, p2, r, unit(dv), s0, s1);
        return (nonzero(n) ? bilinear_density(s0, s1, tri_cosines(p0, p1, p2, r, n)) : 1.0f) / area;
    }
    const float cl = fabsf(dot(ng, dv / __builtin_sqrtf(d2)));
    return cl > 0 ? L.inv_area * d2 / cl : 0.0f;
}

// A light sample from p (normal n, zero in media): selection by the scene's
// light sampler, then the light's proposal; pdf includes the selection pmf
// (for a spot light it is that pmf alone).
RP_DEV LightSample sample_light(const Scene &S, F3 p, F3 n, const Wavelengths &wl, float ul, float u0, float u1) {
    LightSample ls;
    ls.ok = ls.delta = ls.infinite = false;
    ls.pdf = 0;
    ls.dist = 0;
    for (int k = 0; k < 4; ++k) ls.le[k] = 0;
    ls.p = ls.n = ls.err = ls.wi = F3{0, 0, 0};
    const Selected sel = light_select(S, p, n, ul);
    if (sel.light == kEmpty || !(sel.pmf > 0)) return ls;
    const LightEntry &L = S.lights[sel.light];
    if (L.kind == 0 || L.kind == 3) {
        float pdf_omega;
        if (L.kind == 0) {
            F3 a, b, c;
            uint32_t id;
            load_tri(S, L.index, a, b, c, id);
            float su = omin(u0, 0x1.fffffep-1f), sv = omin(u1, 0x1.fffffep-1f), b1, b2;
            const float area = tri_solid_angle(a, b, c, p);
            bool spherical = false;
            if (spherical_tri(area)) {
                // the warp's bilinear density over the solid angle, at the warped sample
                float dens = 1.0f;
                if (nonzero(n)) {
                    const Bilinear w = tri_cosines(a, b, c, p, n);
                    bilinear_warp(su, sv, w);
                    dens = bilinear_density(su, sv, w);
                }
                if (!tri_sample(a, b, c, p, su, sv, b1, b2)) return ls;
                pdf_omega = dens / area;
                spherical = true;
            } else {
                const float root = __builtin_sqrtf(su);
                b1 = 1.0f - root;
                b2 = sv * root;
            }
            const float b0 = 1.0f - b1 - b2;
            ls.p = b0 * a + b1 * b + b2 * c;
            ls.err = gamma_n(6) * (vabs(b0 * a) + vabs(b1 * b) + vabs(b2 * c));
            ls.n = normalize(cross(a - c, b - c));
            if (!spherical) {
                const F3 dv = ls.p - p;
                const float d2 = dot(dv, dv);
                const float cl = fabsf(dot(ls.n, dv / __builtin_sqrtf(d2)));
                pdf_omega = d2 > 0 && cl > 0 ? L.inv_area * d2 / cl : 0.0f;
            }
        } else {
            float su = omin(u0, 0x1.fffffep-1f), sv = omin(u1, 0x1.fffffep-1f), qu = su, qv = sv;
            bool spherical = false;
            if (L.rect) {
                // spherical rectangle above 1e-4 sr: the warp's bilinear
                // density over the solid angle, at the warped sample
                const RectFrame f = rect_frame(L, p);
                if (f.area > 1e-4f) {
                    float dens = 1.0f;
                    if (nonzero(n)) {
                        const Bilinear w = rect_cosines(f, L, n);
                        bilinear_warp(su, sv, w);
                        dens = bilinear_density(su, sv, w);
                    }
                    rect_sample(f, su, sv, qu, qv);
                    pdf_omega = dens / f.area;
                    spherical = true;
                }
            }
            const F3 a = qu * node_vec(L.ex), b = qv * node_vec(L.ey);
            ls.p = node_vec(L.p00) + a + b;
            ls.err = gamma_n(6) * (vabs(node_vec(L.p00)) + vabs(a) + vabs(b));
            ls.n = node_vec(L.uz);
            if (!spherical) pdf_omega = quad_area_pdf(L, p, ls.p);
        }
        const F3 dv = ls.p - p;
        const float d2 = dot(dv, dv);
        if (!(d2 > 0)) return ls;
        ls.dist = __builtin_sqrtf(d2);
        ls.wi = dv / ls.dist;
        const float cos_l = dot(ls.n, -ls.wi);
        const uint32_t obj = __float_as_uint(S.tris[3 * L.index + 1].w);
        const rp_emission &em = S.emissions[S.objects[obj].emission];
        if (!em.two_sided && !(cos_l > 0)) return ls;
        if (cos_l == 0.0f) return ls;
        ls.pdf = sel.pmf * pdf_omega;
        for (int k = 0; k < 4; ++k) ls.le[k] = spectrum_eval(em.radiance, wl.lambda[k], S.spectra);
        ls.ok = ls.pdf > 0 && __builtin_isfinite(ls.pdf);
        return ls;
    }
    if (L.kind == 1) {
        const rp_spot_light &sp = S.spots[L.index];
        F3 pos = {sp.position[0], sp.position[1], sp.position[2]};
        F3 dv = pos - p;
        float d2 = dot(dv, dv);
        if (!(d2 > 0)) return ls;
        ls.dist = __builtin_sqrtf(d2);
        ls.wi = dv / ls.dist;
        F3 axis = {sp.direction[0], sp.direction[1], sp.direction[2]};
        float cos_t = dot(-ls.wi, axis);
        float t = (cos_t - sp.cos_falloff_end) / (sp.cos_falloff_start - sp.cos_falloff_end);
        t = omin(omax(t, 0.0f), 1.0f);
        float fall = t * t * (3.0f - 2.0f * t);
        for (int k = 0; k < 4; ++k) ls.le[k] = spectrum_eval(sp.intensity, wl.lambda[k], S.spectra) * fall / d2;
        ls.p = pos;
        ls.pdf = sel.pmf;
        ls.delta = true;
        ls.ok = true;
        return ls;
    }
    float z = 1.0f - 2.0f * u0;
    float r = __builtin_sqrtf(omax(0.0f, 1.0f - z * z));
    float sn, cs;
    sincos_turns(u1, sn, cs);
    ls.wi = {r * cs, r * sn, z};
    ls.infinite = true;
    ls.pdf = sel.pmf * kInv4Pi;
    for (int k = 0; k < 4; ++k) ls.le[k] = spectrum_eval(S.environment, wl.lambda[k], S.spectra);
    ls.ok = true;
    return ls;
}

// Solid-angle density of light sampling from (p, n) toward emitter point q
// (geometric normal ng) of light `id`: selection pmf x proposal density.
RP_DEV float area_light_pdf(const Scene &S, uint32_t id, F3 p, F3 n, F3 q, F3 ng) {
    const float pmf = light_pmf(S, p, n, id);
    const LightEntry &L = S.lights[id];
    if (L.kind == 3) {
        float u, v;
        quad_coords(L, q, u, v);
        return pmf * quad_pdf(L, p, n, u, v, q);
    }
    F3 a, b, c;
    uint32_t tri_id;
    load_tri(S, L.index, a, b, c, tri_id);
    return pmf * tri_pdf(L, a, b, c, p, n, q, ng);
}

// Visibility with transmittance through null interfaces and media.
// With `mlen` the length travelled inside media is accumulated (adjoint).
// kInterfaces = false compiles only the any-hit branch (scenes without null
// interfaces), which keeps the shadow kernel's register footprint small.
// With kCount the traversals' node fetches and triangle tests add to `tc`.
template <bool kInterfaces = true, bool kCount = false>
RP_DEV bool transmittance(const Scene &S, Stack &st, F3 o, F3 d, float tmax, uint32_t medium, const Wavelengths &wl,
                          float tr[4], uint32_t boundary_limit, float *mlen = nullptr,
                          uint32_t mlen_medium = RP_NONE, TraversalCounts *tc = nullptr) {
    for (int k = 0; k < 4; ++k) tr[k] = 1.0f;
    if (mlen) *mlen = 0.0f;
    if (!kInterfaces || !S.has_interfaces) {
        Hit h;
        if (traverse<true, kCount>(S, st, o, d, tmax, h, tc)) return false;
        if (medium != RP_NONE) {
            MediumCoeff mc = medium_coeff(S, medium, wl);
            float len = length(d);
            for (int k = 0; k < 4; ++k) tr[k] = tr[k] * exp_f(-mc.st[k] * (tmax * len));
            if (mlen && medium == mlen_medium) *mlen = *mlen + tmax * len;
        }
        return true;
    }
    uint32_t crossings = 0;
    for (;;) {
        Hit h;
        bool hit = traverse<false, kCount>(S, st, o, d, tmax, h, tc);
        float seg = hit ? h.t : tmax;
        if (medium != RP_NONE) {
            MediumCoeff mc = medium_coeff(S, medium, wl);
            float len = length(d);
            for (int k = 0; k < 4; ++k) tr[k] = tr[k] * exp_f(-mc.st[k] * (seg * len));
            if (mlen && medium == mlen_medium) *mlen = *mlen + seg * len;
        }
        if (!hit) return true;
        uint32_t obj = __float_as_uint(S.tris[3 * h.prim + 1].w);
        const rp_object &ob = S.objects[obj];
        if (S.materials[ob.material].kind != RP_MATERIAL_INTERFACE) return false;
        if (++crossings > boundary_limit) return false;
        SurfacePoint sp = surface_point(S, h);
        medium = outgoing_medium(ob, medium, sp.n, d);
        F3 no = offset_origin(sp.p, sp.err, sp.n, d);
        F3 target = o + tmax * d;
        if (__builtin_isinf(tmax)) {
            o = no;
        } else {
            d = target - no;
            o = no;
            tmax = 1.0f;
        }
    }
}

// ======================================================== hero bookkeeping
RP_DEV float mixture(const float d[4], const float p[4]) { return (d[0] * p[0] + d[1] * p[1]) + (d[2] * p[2] + d[3] * p[3]); }
RP_DEV float dsum(const float d[4]) { return (d[0] + d[1]) + (d[2] + d[3]); }
RP_DEV void normalize_density(float d[4]) {
    float m = omax(omax(d[0], d[1]), omax(d[2], d[3]));
    if (!(m > 0) || !__builtin_isfinite(m)) return;
    int e;
    (void)frexpf(m, &e);
    for (int k = 0; k < 4; ++k) d[k] = ldexpf(d[k], -e);
}
RP_DEV bool advance(float density[4], float beta[4], const float num[4], const float pdf[4]) {
    float nd[4];
    for (int k = 0; k < 4; ++k) nd[k] = density[k] * pdf[k];
    float before = dsum(density), after = dsum(nd);
    if (!(before > 0 && after > 0) || !__builtin_isfinite(after)) return false;
    float ratio = before / after;
    for (int k = 0; k < 4; ++k) {
        beta[k] = (beta[k] * num[k]) * ratio;
        if (!__builtin_isfinite(beta[k])) return false;
        density[k] = nd[k];
    }
    normalize_density(density);
    return true;
}
RP_DEV bool positive_normal(float v) {
    uint32_t w = __float_as_uint(v);
    return w >= 0x00800000u && w < 0x7f800000u;
}
RP_DEV bool hero_continue(float beta[4], const float eta_scale[4], bool enabled, uint32_t depth, float u) {
    if (!enabled || depth <= 1) return true;
    float maximum = 0.0f;
    for (int k = 0; k < 4; ++k) {
        float v = beta[k] * eta_scale[k];
        if (!__builtin_isfinite(v) || v < 0 || (beta[k] > 0 && !positive_normal(v))) return true;
        maximum = omax(maximum, v);
    }
    if (!positive_normal(maximum) || maximum >= 1.0f) return true;
    float q = 1.0f - maximum;
    float survival = 1.0f - q;
    if (!(q > 0 && q < 1) || !positive_normal(survival)) return true;
    float surv[4];
    for (int k = 0; k < 4; ++k) {
        surv[k] = beta[k];
        if (__float_as_uint(beta[k]) == 0) continue;
        surv[k] = beta[k] / survival;
        if (!positive_normal(surv[k])) return true;
    }
    if (u < q) return false;
    for (int k = 0; k < 4; ++k) beta[k] = surv[k];
    return true;
}

// Hero sensor: sum_k cie(lambda_k) * (L_k / pdf_k / 4).
RP_DEV void sensor_xyz(const Scene &S, const Wavelengths &wl, const float L[4], float xyz[3]) {
    xyz[0] = xyz[1] = xyz[2] = 0.0f;
    for (int k = 0; k < 4; ++k) {
        if (!(wl.pdf[k] > 0)) continue;
        int i = int(floorf(wl.lambda[k] + 0.5f)) - 360;
        i = i < 0 ? 0 : (i > 470 ? 470 : i);
        float w = (L[k] / wl.pdf[k]) / 4.0f;
        float4 c = S.cie[i];
        xyz[0] = xyz[0] + c.x * w;
        xyz[1] = xyz[1] + c.y * w;
        xyz[2] = xyz[2] + c.z * w;
    }
}

} // namespace rpd
