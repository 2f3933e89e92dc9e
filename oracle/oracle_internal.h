// Oracle: CPU restatement of the reference transport for parity testing.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// or as the timed CPU baseline -- never as the product path.
//
// Arithmetic contract shared with the HIP kernels (so that per-path results
// agree bit for bit): binary32, no FMA contraction (-ffp-contract=off; fma only
// where written explicitly), IEEE division and square root, and the
// transcendental approximations defined below (no libm on the path).
#pragma once
#include "../include/rp_api.h"

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cfloat>
#include <cstring>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- vectors
struct V3 {
    float x = 0, y = 0, z = 0;
};
inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline float get(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// Kahan's difference of products a*b - c*d (PBRT v4 DifferenceOfProducts).
inline float dop(float a, float b, float c, float d) {
    float cd = c * d;
    float err = std::fma(-c, d, cd);
    float r = std::fma(a, b, -cd);
    return r + err;
}
inline V3 cross(V3 a, V3 b) {
    return {dop(a.y, b.z, a.z, b.y), dop(a.z, b.x, a.x, b.z), dop(a.x, b.y, a.y, b.x)};
}
// Plain cross product (the light-sampling geometry uses this form).
inline V3 cross3(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv4Pi = 0.07957747154594766788f;
inline float gamma_n(int n) { return (float(n) * 0x1p-24f) / (1.0f - float(n) * 0x1p-24f); }

// ------------------------------------------------------ transcendentals
// sin/cos of 2*pi*t: quadrant reduction in turns (exact), then truncated
// Taylor series on |a| <= pi/4 evaluated by Horner's rule.
inline void sincos_turns(float t, float &s, float &c) {
    float x = t * 4.0f;
    float q = std::floor(x + 0.5f);
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
// exp: Cody-Waite reduction x = k ln2 + r (two-part ln2), degree-7 Taylor.
inline float exp_f(float x) {
    if (x != x) return x;
    if (!(x > -104.0f)) return 0.0f;
    if (x > 88.72f) return INFINITY;
    float k = std::floor(x * 1.44269504088896341f + 0.5f);
    float r = (x - k * 0.693145751953125f) - k * 1.42860682030941723212e-6f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (1.0f / 6.0f + r * (1.0f / 24.0f + r * (1.0f / 120.0f + r * (1.0f / 720.0f + r * (1.0f / 5040.0f)))))));
    return std::ldexp(p, int(k));
}
// log: frexp to m in [sqrt(1/2), sqrt(2)), atanh series in s = (m-1)/(m+1).
inline float log_f(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e;
    float m = std::frexp(x, &e);
    if (m < 0.70710678118654752f) { m = m * 2.0f; e -= 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float lm = s * (2.0f + s2 * (2.0f / 3.0f + s2 * (2.0f / 5.0f + s2 * (2.0f / 7.0f + s2 * (2.0f / 9.0f)))));
    float ef = float(e);
    return ef * 0.693145751953125f + (lm + ef * 1.42860682030941723212e-6f);
}
// sin and cos of an angle in radians through the turn reduction above.
inline float sin_f(float x) {
    float s, c;
    sincos_turns(x * 0.159154943091895336f, s, c);
    return s;
}
inline float cos_f(float x) {
    float s, c;
    sincos_turns(x * 0.159154943091895336f, s, c);
    return c;
}
// atan: |x| > 1 through pi/2 - atan(1/|x|), then z > tan(pi/8) through
// pi/4 + atan((z-1)/(z+1)), leaving |t| <= tan(pi/8) for the Maclaurin
// series t - t^3/3 + ... - t^15/15 + t^17/17 (truncation below 3e-9).
inline float atan_f(float x) {
    float ax = std::fabs(x);
    bool inv = ax > 1.0f;
    float z = inv ? 1.0f / ax : ax;
    bool shift = z > 0.414213562373095049f;
    float t = shift ? (z - 1.0f) / (z + 1.0f) : z;
    float t2 = t * t;
    float p = t + t * t2 * (-1.0f / 3.0f + t2 * (1.0f / 5.0f + t2 * (-1.0f / 7.0f + t2 * (1.0f / 9.0f +
              t2 * (-1.0f / 11.0f + t2 * (1.0f / 13.0f + t2 * (-1.0f / 15.0f + t2 * (1.0f / 17.0f))))))));
    float r = shift ? 0.785398163397448310f + p : p;
    r = inv ? 1.57079632679489662f - r : r;
    return std::copysign(r, x);
}
inline float atan2_f(float y, float x) {
    float ax = std::fabs(x), ay = std::fabs(y);
    if (ax == 0.0f && ay == 0.0f) return std::signbit(x) ? std::copysign(3.14159265358979324f, y) : std::copysign(0.0f, y);
    float r = ay <= ax ? atan_f(ay / ax) : 1.57079632679489662f - atan_f(ax / ay);
    if (std::signbit(x)) r = 3.14159265358979324f - r;
    return std::copysign(r, y);
}
inline float asin_f(float x) { return atan2_f(x, std::sqrt(std::max(0.0f, (1.0f - x) * (1.0f + x)))); }

// Light-sampling context normal: a unit normal through a 2 x 16-bit
// octahedral code (snorm, round half up) and back, so that the normal a
// vertex samples lights with and the one the next vertex's MIS weight uses
// are the same bits.  Code kNoNormal (media, the camera) decodes to zero.
constexpr uint32_t kNoNormal = 0xffffffffu;
inline float sign_nz(float v) { return v < 0.0f ? -1.0f : 1.0f; }
inline uint32_t oct_encode(V3 n) {
    const float l1 = std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z);
    float px = n.x / l1, py = n.y / l1;
    if (n.z < 0.0f) {
        const float qx = (1.0f - std::fabs(py)) * sign_nz(px), qy = (1.0f - std::fabs(px)) * sign_nz(py);
        px = qx;
        py = qy;
    }
    const int32_t ix = int32_t(std::floor(std::clamp(px, -1.0f, 1.0f) * 32767.0f + 0.5f));
    const int32_t iy = int32_t(std::floor(std::clamp(py, -1.0f, 1.0f) * 32767.0f + 0.5f));
    return (uint32_t(ix) & 0xffffu) | (uint32_t(iy) << 16);
}
inline V3 oct_decode(uint32_t c) {
    if (c == kNoNormal) return {0, 0, 0};
    float px = float(int16_t(c & 0xffffu)) / 32767.0f, py = float(int16_t(c >> 16)) / 32767.0f;
    const float z = 1.0f - std::fabs(px) - std::fabs(py);
    if (z < 0.0f) {
        const float qx = (1.0f - std::fabs(py)) * sign_nz(px), qy = (1.0f - std::fabs(px)) * sign_nz(py);
        px = qx;
        py = qy;
    }
    const V3 v = {px, py, z};
    return v / std::sqrt(dot(v, v));
}

inline float next_up(float v) {
    if (std::isinf(v) && v > 0) return v;
    if (v == -0.0f) v = 0.0f;
    uint32_t u = std::bit_cast<uint32_t>(v);
    if (v >= 0) ++u; else --u;
    return std::bit_cast<float>(u);
}
inline float next_down(float v) {
    if (std::isinf(v) && v < 0) return v;
    if (v == 0.0f) v = -0.0f;
    uint32_t u = std::bit_cast<uint32_t>(v);
    if (v > 0) --u; else ++u;
    return std::bit_cast<float>(u);
}

// ---------------------------------------------------------------- sampling
// Sobol tables (PBRT v4 layout: directions[1024*52] u32, intervals[25*52] u64,
// inverse[26*52] u64); restated from src/sobol_shared.h:30-41.
struct SobolTables {
    std::vector<uint32_t> directions;
    std::vector<uint64_t> intervals;
    std::vector<uint64_t> inverse;
};
const SobolTables &sobol_tables();
bool load_sobol_tables(const char *path);

uint64_t sobol_index(uint32_t m, uint32_t x, uint32_t y, uint64_t sample);
uint32_t sobol_dimension_hash(uint32_t dimension, uint32_t seed_bits);
uint32_t sobol_word(uint64_t index, uint32_t dimension, uint32_t mode, uint32_t seed_bits);
float sobol_uniform(uint32_t word);
float sobol_camera_uniform(uint32_t word, uint32_t m, uint32_t pixel);
std::array<uint32_t, 4> philox(std::array<uint32_t, 4> counter, std::array<uint32_t, 2> key);
float philox_u01(uint32_t word);
uint32_t sobol_scale_exponent(uint32_t width, uint32_t height);

struct Wavelengths {
    float lambda[4];
    float pdf[4];
};
Wavelengths sample_wavelengths(float u, uint32_t mode);

// ---------------------------------------------------------------- spectra
struct CieTable {
    float xyz[471][3];
};
const CieTable &cie_table();
bool load_cie_table(const char *csv_path);
float spectrum_eval(const rp_spectrum &s, float lambda, const float *pool);

} // namespace orc
