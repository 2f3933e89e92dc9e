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
