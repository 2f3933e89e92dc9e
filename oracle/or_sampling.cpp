// Oracle sampling: Sobol (PBRT v4 interval index + four randomizers), Philox
// 4x32-10, hero wavelength sampling, CIE sensor table and spectrum evaluation.
// TEST INFRASTRUCTURE ONLY (see oracle_internal.h).
#include "oracle_internal.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

namespace orc {

namespace {
SobolTables g_sobol;
CieTable g_cie;
} // namespace

const SobolTables &sobol_tables() { return g_sobol; }
const CieTable &cie_table() { return g_cie; }

bool load_sobol_tables(const char *path) {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    g_sobol.directions.resize(1024 * 52);
    g_sobol.intervals.resize(25 * 52);
    g_sobol.inverse.resize(26 * 52);
    bool ok = std::fread(g_sobol.directions.data(), 4, g_sobol.directions.size(), f) == g_sobol.directions.size() &&
              std::fread(g_sobol.intervals.data(), 8, g_sobol.intervals.size(), f) == g_sobol.intervals.size() &&
              std::fread(g_sobol.inverse.data(), 8, g_sobol.inverse.size(), f) == g_sobol.inverse.size();
    std::fclose(f);
    return ok;
}

bool load_cie_table(const char *csv_path) {
    std::ifstream in(csv_path);
    if (!in) return false;
    std::string line;
    int row = 0;
    while (std::getline(in, line) && row < 471) {
        if (line.empty()) continue;
        int nm = 0;
        double x = 0, y = 0, z = 0;
        if (std::sscanf(line.c_str(), "%d,%lf,%lf,%lf", &nm, &x, &y, &z) != 4) return false;
        if (nm != 360 + row) return false;
        g_cie.xyz[row][0] = float(x);
        g_cie.xyz[row][1] = float(y);
        g_cie.xyz[row][2] = float(z);
        ++row;
    }
    return row == 471;
}

// PBRT interval index: src/sobol_shared.h:56-76 (lowdiscrepancy.h SobolIntervalToIndex).
uint64_t sobol_index(uint32_t m, uint32_t x, uint32_t y, uint64_t sample) {
    if (m == 0) return sample;
    const auto &t = g_sobol;
    uint32_t row = (m - 1) * 52;
    uint64_t delta = 0;
    uint64_t result = sample << (2 * m);
    uint32_t col = row;
    for (uint64_t s = sample; s; s >>= 1, ++col)
        if (s & 1) delta ^= t.intervals[col];
    uint64_t bits = ((uint64_t(x) << m) | y) ^ delta;
    col = row;
    for (; bits; bits >>= 1, ++col)
        if (bits & 1) result ^= t.inverse[col];
    return result;
}

// MurmurHash64A over LE32(dimension) || LE32(seed): src/sobol_shared.h:78-89.
uint32_t sobol_dimension_hash(uint32_t dimension, uint32_t seed_bits) {
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    uint64_t k = ((uint64_t(seed_bits) << 32) | dimension) * m;
    k ^= k >> 47;
    k *= m;
    uint64_t h = (8 * m) ^ k;
    h *= m;
    h ^= h >> 47;
    h *= m;
    h ^= h >> 47;
    return uint32_t(h);
}

static uint32_t reverse32(uint32_t v) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((v >> i) & 1u) << (31 - i);
    return r;
}
static uint64_t mix64(uint64_t v) {
    v ^= v >> 31;
    v *= 0x7fb5d329728ea185ull;
    v ^= v >> 27;
    v *= 0x81dadef4bc2dd44dull;
    return v ^ (v >> 33);
}

// src/sobol_shared.h:108-141: direction XOR then none/permute/fastowen/owen.
uint32_t sobol_word(uint64_t index, uint32_t dimension, uint32_t mode, uint32_t seed_bits) {
    const auto &t = g_sobol;
    uint32_t v = 0;
    uint32_t col = dimension * 52;
    for (uint64_t i = index; i; i >>= 1, ++col)
        if (i & 1) v ^= t.directions[col];
    if (mode == RP_RANDOMIZE_NONE) return v;
    uint32_t hash = sobol_dimension_hash(dimension, seed_bits);
    if (mode == RP_RANDOMIZE_PERMUTE) return v ^ hash;
    if (mode == RP_RANDOMIZE_FASTOWEN) {
        v = reverse32(v);
        v ^= v * 0x3d20adeau;
        v += hash;
        v *= (hash >> 16) | 1u;
        v ^= v * 0x05526c56u;
        v ^= v * 0x53a22864u;
        return reverse32(v);
    }
    v ^= (hash & 1u) << 31;
    for (uint32_t b = 1; b < 32; ++b) {
        uint32_t prefix = v & (0xffffffffu << (32 - b));
        if (uint32_t(mix64(uint64_t(prefix ^ hash))) & (1u << b)) v ^= 1u << (31 - b);
    }
    return v;
}

// src/sobol_shared.h:143-154.
float sobol_uniform(uint32_t word) {
    float u = float(word) * 0x1p-32f;
    return u < 0x1.fffffep-1f ? u : 0x1.fffffep-1f;
}
float sobol_camera_uniform(uint32_t word, uint32_t m, uint32_t pixel) {
    float local = sobol_uniform(word) * float(uint32_t(1) << m) - float(pixel);
    if (local < 0.0f) return 0.0f;
    return local < 0x1.fffffep-1f ? local : 0x1.fffffep-1f;
}

uint32_t sobol_scale_exponent(uint32_t width, uint32_t height) {
    uint32_t v = std::max(std::max(width, height), 1u) - 1;
    uint32_t m = 0;
    while (v) { ++m; v >>= 1; }
    return m;
}

// Philox4x32-10 (Salmon et al. 2011), shaders/render.metal:97-109.
std::array<uint32_t, 4> philox(std::array<uint32_t, 4> c, std::array<uint32_t, 2> k) {
    const uint32_t m0 = 0xd2511f53u, m1 = 0xcd9e8d57u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = uint64_t(m0) * c[0];
        uint64_t p1 = uint64_t(m1) * c[2];
        std::array<uint32_t, 4> n = {uint32_t(p1 >> 32) ^ c[1] ^ k[0], uint32_t(p1),
                                     uint32_t(p0 >> 32) ^ c[3] ^ k[1], uint32_t(p0)};
        c = n;
        k[0] += 0x9e3779b9u;
        k[1] += 0xbb67ae85u;
    }
    return c;
}
float philox_u01(uint32_t w) { return float(w >> 8) * 0x1p-24f; }

// Hero wavelengths, src/hero_spectrum_shared.h:199-244.
Wavelengths sample_wavelengths(float u, uint32_t mode) {
    Wavelengths w{};
    if (!(u >= 0 && u < 1)) {
        for (int k = 0; k < 4; ++k) w.lambda[k] = w.pdf[k] = NAN;
        return w;
    }
    if (mode == 0) {
        for (int k = 0; k < 4; ++k) {
            float s = u + float(k) * 0.25f;
            if (s >= 1.0f) s -= 1.0f;
            float l = 360.0f + 470.0f * s;
            if (l >= 830.0f) l = std::bit_cast<float>(std::bit_cast<uint32_t>(830.0f) - 1);
            w.lambda[k] = l;
            w.pdf[k] = 1.0f / 470.0f;
        }
        return w;
    }
    // PBRT visible-wavelength fit on integer-shifted 24-bit lattice coordinates.
    for (int k = 0; k < 4; ++k) {
        uint32_t coord = uint32_t(u * 16777216.0f);
        float s = float((coord + uint32_t(k) * 4194304u) & 0x00ffffffu) * 0x1p-24f;
        float a = 0.85691062f - 1.82750197f * s;
        float at = 0.5f * log_f((1.0f + a) / (1.0f - a));
        float l = 538.0f - 138.888889f * at;
        l = std::min(std::max(l, 360.0f), 830.0f);
        float e = exp_f(0.0072f * (l - 538.0f));
        float ch = 0.5f * (e + 1.0f / e);
        w.lambda[k] = l;
        w.pdf[k] = 0.0039398042f / (ch * ch);
    }
    return w;
}

// Spectrum evaluation at one wavelength (nanometres).  Tables use the
// nearest integer nanometre with halves rounded up (shaders/frame_sensor.h:10-20).
float spectrum_eval(const rp_spectrum &s, float l, const float *pool) {
    float v = 0.0f;
    switch (s.kind) {
    case RP_SPECTRUM_CONSTANT: v = s.c[0]; break;
    case RP_SPECTRUM_SIGMOID: {
        float x = std::fma(std::fma(s.c[0], l, s.c[1]), l, s.c[2]);
        if (std::isinf(x)) v = x > 0 ? 1.0f : 0.0f;
        else v = 0.5f + x / (2.0f * std::sqrt(1.0f + x * x));
        break;
    }
    case RP_SPECTRUM_TABLE: {
        int i = int(std::floor(l + 0.5f)) - 360;
        i = std::min(std::max(i, 0), 470);
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
    default: v = NAN;
    }
    return v * s.scale;
}

} // namespace orc
