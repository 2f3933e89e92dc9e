// Host services of the rypath_amd library: errors, data tables, spectra,
// light tables and scene validation.
#include "rp_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <functional>
#include <mutex>
#include <stdexcept>

namespace rp {

namespace {
thread_local std::string g_error;
}
void set_error(const std::string &message) { g_error = message; }
const std::string &last_error() { return g_error; }

std::string data_dir() {
    if (const char *env = std::getenv("RYPATH_AMD_DATA")) return env;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void *>(&rp_api_version), &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        auto slash = p.find_last_of('/');
        return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/data";
    }
    return "data";
}

const SobolData &sobol_data() {
    static SobolData data;
    static std::once_flag once;
    std::call_once(once, [] {
        std::string path = data_dir() + "/sobol_tables.bin";
        std::FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + path);
        data.directions.resize(1024 * 52);
        data.intervals.resize(25 * 52);
        data.inverse.resize(26 * 52);
        bool ok = std::fread(data.directions.data(), 4, data.directions.size(), f) == data.directions.size() &&
                  std::fread(data.intervals.data(), 8, data.intervals.size(), f) == data.intervals.size() &&
                  std::fread(data.inverse.data(), 8, data.inverse.size(), f) == data.inverse.size();
        std::fclose(f);
        if (!ok) throw std::runtime_error("short read " + path);
    });
    return data;
}

const std::array<std::array<float, 3>, 471> &cie_xyz() {
    static std::array<std::array<float, 3>, 471> table{};
    static std::once_flag once;
    std::call_once(once, [] {
        std::string path = data_dir() + "/cie_xyz_1931_2deg.csv";
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot open " + path);
        std::string line;
        int row = 0;
        while (std::getline(in, line) && row < 471) {
            if (line.empty()) continue;
            int nm = 0;
            double x, y, z;
            if (std::sscanf(line.c_str(), "%d,%lf,%lf,%lf", &nm, &x, &y, &z) != 4 || nm != 360 + row)
                throw std::runtime_error("bad CIE row in " + path);
            table[row] = {float(x), float(y), float(z)};
            ++row;
        }
        if (row != 471) throw std::runtime_error("short CIE table " + path);
    });
    return table;
}

const std::array<float, 471> &d65() {
    static std::array<float, 471> table{};
    static std::once_flag once;
    std::call_once(once, [] {
        std::string path = data_dir() + "/d65.bin";
        std::FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + path);
        bool ok = std::fread(table.data(), 4, 471, f) == 471;
        std::fclose(f);
        if (!ok) throw std::runtime_error("short read " + path);
    });
    return table;
}

double y_inner_product(const std::array<float, 471> &s) {
    const auto &cie = cie_xyz();
    double sum = 0;
    for (size_t i = 0; i < 471; ++i) sum += double(s[i]) * double(cie[i][1]);
    return sum;
}

// ------------------------------------------------------------- sRGB sigmoid
namespace {
struct SigmoidTable {
    std::vector<float> z_nodes;  // 64
    std::vector<float> coeffs;   // [3][64][64][64][3]
};
const SigmoidTable &sigmoid_table() {
    static SigmoidTable t;
    static std::once_flag once;
    std::call_once(once, [] {
        std::string path = data_dir() + "/srgb_sigmoid.bin";
        std::FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + path);
        t.z_nodes.resize(64);
        t.coeffs.resize(size_t(3) * 64 * 64 * 64 * 3);
        bool ok = std::fread(t.z_nodes.data(), 4, 64, f) == 64 &&
                  std::fread(t.coeffs.data(), 4, t.coeffs.size(), f) == t.coeffs.size();
        std::fclose(f);
        if (!ok) throw std::runtime_error("short read " + path);
    });
    return t;
}
float lerpf(float t, float a, float b) { return (1.0f - t) * a + t * b; }
} // namespace

// PBRT v4 RGBToSpectrumTable::operator() restated (trilinear lookup of the
// Jakob-Hanika coefficient table).
SigmoidCoefficients srgb_sigmoid(const std::array<float, 3> &rgb) {
    SigmoidCoefficients out;
    if (rgb[0] == rgb[1] && rgb[1] == rgb[2]) {
        out.constant = true;
        out.value = rgb[0];
        return out;
    }
    const auto &t = sigmoid_table();
    const int res = 64;
    int maxc = 0;
    for (int ch = 1; ch < 3; ++ch)
        if (rgb[ch] >= rgb[maxc]) maxc = ch;
    float z = rgb[maxc];
    float x = rgb[(maxc + 1) % 3] * float(res - 1) / z;
    float y = rgb[(maxc + 2) % 3] * float(res - 1) / z;
    // FindInterval over the z nodes
    int lo = 0, size = res - 2, first = 1;
    (void)lo;
    {
        int sz = res - 2;
        first = 1;
        while (sz > 0) {
            int half = sz >> 1, middle = first + half;
            if (t.z_nodes[middle] < z) {
                first = middle + 1;
                sz -= half + 1;
            } else sz = half;
        }
    }
    (void)size;
    int zi = std::min(std::max(first - 1, 0), res - 2);
    float dz = (z - t.z_nodes[zi]) / (t.z_nodes[zi + 1] - t.z_nodes[zi]);
    int xi = std::min(int(x), res - 2), yi = std::min(int(y), res - 2);
    float dx = x - float(xi), dy = y - float(yi);
    auto co = [&](int ox, int oy, int oz, int i) {
        size_t idx = ((((size_t(maxc) * res + size_t(zi + oz)) * res + size_t(yi + oy)) * res + size_t(xi + ox)) * 3) + i;
        return t.coeffs[idx];
    };
    for (int i = 0; i < 3; ++i) {
        out.c[i] = lerpf(dz, lerpf(dy, lerpf(dx, co(0, 0, 0, i), co(1, 0, 0, i)), lerpf(dx, co(0, 1, 0, i), co(1, 1, 0, i))),
                         lerpf(dy, lerpf(dx, co(0, 0, 1, i), co(1, 0, 1, i)), lerpf(dx, co(0, 1, 1, i), co(1, 1, 1, i))));
    }
    return out;
}

rp_spectrum spectrum_from_rgb(const std::array<float, 3> &rgb) {
    SigmoidCoefficients c = srgb_sigmoid(rgb);
    rp_spectrum s{};
    s.scale = 1.0f;
    if (c.constant) {
        s.kind = RP_SPECTRUM_CONSTANT;
        s.c[0] = c.value;
    } else {
        s.kind = RP_SPECTRUM_SIGMOID;
        s.c[0] = c.c[0];
        s.c[1] = c.c[1];
        s.c[2] = c.c[2];
    }
    return s;
}

// Planck's law normalised to the Wien peak (c2 = 1.4387769e7 nm K).
std::array<float, 471> blackbody_table(double kelvin, double scale) {
    std::array<float, 471> out{};
    const double c = 299792458.0, h = 6.62606957e-34, kb = 1.3806488e-23;
    auto planck = [&](double nm) {
        double l = nm * 1e-9;
        return (2 * h * c * c) / (std::pow(l, 5) * (std::exp((h * c) / (l * kb * kelvin)) - 1));
    };
    double peak_nm = 2.8977721e-3 / kelvin * 1e9;
    double peak = planck(peak_nm);
    for (int i = 0; i < 471; ++i) out[i] = float(scale * planck(360.0 + i) / peak);
    return out;
}

float spectrum_eval_host(const rp_spectrum &s, float l, const float *pool) {
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

// ------------------------------------------------------------- lights
float scene_bounding_radius(const rp_scene_desc &d) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = 0; v < d.vertex_count; ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], double(d.positions[3 * v + a]));
            hi[a] = std::max(hi[a], double(d.positions[3 * v + a]));
        }
    double r2 = 0;
    for (int a = 0; a < 3; ++a) {
        double h = 0.5 * (hi[a] - lo[a]);
        r2 += h * h;
    }
    return d.vertex_count ? float(std::sqrt(r2)) : 0.0f;
}

// Light tables for next-event estimation (the reference's preparation,
// src/metal_render.mm:540-640 and 1560-1780, and build_light_sampling,
// src/light_sampling.cpp:226-571; PBRT v4's BVHLightSampler and
// PowerLightSampler): one light per emissive triangle, per quad (two
// consecutive triangles (a,b,c), (a,c,d) of one emissive object forming a
// rectangle), per spot and for the environment; their power alias table;
// the light BVH over the finite lights.  Double precision, in the order the
// reference builds them.
namespace {

// Power proxy of a spectrum (light_spectrum_power, src/light_sampling.cpp:17-27):
// the visible-wavelength estimate at u = 1/2 of its integral over nanometres,
// at the packet sampler's four wavelengths (float shifts by 1/4, a shifted
// value above 1 wrapped; src/packet_spectrum_shared.h:27-58) with the host's
// binary32 atanh and cosh, as the reference's host computes it; zero: the
// plain sum over 360..830 nm.
double spectrum_power(const rp_spectrum &sp, const float *pool) {
    double v = 0;
    for (int k = 0; k < 4; ++k) {
        float up = 0.5f + float(k) * 0.25f;
        if (up > 1) up -= 1;
        float l = 538 - 138.888889f * std::atanh(0.85691062f - 1.82750197f * up);
        l = std::min(std::max(l, 360.0f), 830.0f);
        const float c = std::cosh(0.0072f * (l - 538));
        const float pdf = 0.0039398042f / (c * c);
        v += double(spectrum_eval_host(sp, std::round(l), pool)) / pdf / 4;
    }
    if (v == 0)
        for (int i = 0; i < 471; ++i) v += double(spectrum_eval_host(sp, float(360 + i), pool));
    return v;
}
double spectrum_max(const rp_spectrum &sp, const float *pool) {
    double m = 0;
    for (int i = 0; i < 471; ++i) m = std::max(m, double(spectrum_eval_host(sp, float(360 + i), pool)));
    return m;
}

struct Bounds {
    double power = 0, bounds_power = -1;
    float low[3] = {0, 0, 0}, high[3] = {0, 0, 0}, dir[3] = {0, 0, 1};
    float cos_outer = -1, cos_emission = 0;
    bool two_sided = true, infinite = false;
};
using D3 = std::array<double, 3>;
D3 dsub(const D3 &a, const D3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
double ddot(const D3 &a, const D3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double dlen(const D3 &a) { return std::sqrt(ddot(a, a)); }
D3 dcross(const D3 &a, const D3 &b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// Union: boxes, powers, two-sidedness, the smaller emission cosine, and
// the smallest direction cone holding both cones.
Bounds unite(const Bounds &a, const Bounds &b) {
    if (a.power == 0) return b;
    if (b.power == 0) return a;
    Bounds r = a;
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
    if (std::min(between + tb, pi) <= ta) return r;
    if (std::min(between + ta, pi) <= tb) {
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
    D3 t;
    for (int i = 0; i < 3; ++i) t[i] = wb[i] - ddot(wa, wb) * wa[i];
    double tl = dlen(t);
    if (tl < 1e-12) {
        const int axis = std::fabs(wa[0]) < 0.8 ? 0 : 1;
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
// Split cost: power x the cone's angular measure x surface area, scaled by
// the parent's aspect ratio along `axis`.
double split_cost(const Bounds &b, const Bounds &all, int axis) {
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
// Rectangle classification of p = (p00, p10, p01, p11) in double
// (src/light_sampling.cpp:294-311) and the sampler's binary32 test.
bool rectangle_f64(const D3 p[4]) {
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
struct F3h {
    float x, y, z;
};
F3h fsub(F3h a, F3h b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3h fadd(F3h a, F3h b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
float fdot(F3h a, F3h b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
bool rectangle_f32(F3h p00, F3h p10, F3h p01, F3h p11) {
    const F3h ex = fsub(p10, p00), ey = fsub(p01, p00);
    const float x2 = fdot(ex, ex), y2 = fdot(ey, ey);
    if (!(x2 > 0 && y2 > 0)) return false;
    const F3h e = fsub(p11, fadd(fadd(p00, ex), ey));
    return fdot(e, e) <= 0x1p-40f * std::max(x2, y2) && std::fabs(fdot(ex, ey)) <= 0x1p-20f * std::sqrt(x2 * y2);
}

// The smallest-angle cone about the mean unit normal holding n normals
// (light_normal_cone, src/light_sampling.cpp:356-381): a degenerate normal or
// a vanishing mean leaves the full sphere (cos_outer = -1).
void normal_cone(Bounds &b, const D3 *ns, int n) {
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

} // namespace

LightTables build_light_tables(const rp_scene_desc &d, float scene_radius, std::vector<uint32_t> &tri_light) {
    const double pi = 3.14159265358979323846;
    LightTables out;
    tri_light.assign(d.triangle_count, RP_NONE);
    std::vector<Bounds> in;
    auto vert = [&](uint32_t i) {
        return F3h{d.positions[3 * size_t(i)], d.positions[3 * size_t(i) + 1], d.positions[3 * size_t(i) + 2]};
    };
    auto dv = [](F3h v) { return D3{double(v.x), double(v.y), double(v.z)}; };
    auto box_of = [](Bounds &b, const F3h *c, int n) {
        for (int a = 0; a < 3; ++a) {
            b.low[a] = INFINITY;
            b.high[a] = -INFINITY;
        }
        for (int k = 0; k < n; ++k) {
            const float v[3] = {c[k].x, c[k].y, c[k].z};
            for (int a = 0; a < 3; ++a) {
                b.low[a] = std::min(b.low[a], v[a]);
                b.high[a] = std::max(b.high[a], v[a]);
            }
        }
    };
    for (uint32_t t = 0; t < d.triangle_count; ++t) {
        const uint32_t o = d.tri_object[t];
        const rp_object &obj = d.objects[o];
        if (obj.emission == RP_NONE) continue;
        const rp_emission &em = d.emissions[obj.emission];
        const double mean = spectrum_power(em.radiance, d.spectra), bound = spectrum_max(em.radiance, d.spectra);
        const uint32_t *ia = d.indices + 3 * size_t(t);
        if (t + 1 < d.triangle_count && d.tri_object[t + 1] == o && ia[3] == ia[0] && ia[4] == ia[2]) {
            const F3h c[4] = {vert(ia[0]), vert(ia[1]), vert(ia[5]), vert(ia[2])}; // p00, p10, p01, p11
            const D3 p[4] = {dv(c[0]), dv(c[1]), dv(c[2]), dv(c[3])};
            if (rectangle_f64(p) && rectangle_f32(c[0], c[1], c[2], c[3])) {
                const double area = dlen(dsub(p[2], p[0])) * dlen(dsub(p[1], p[0]));
                Bounds b;
                box_of(b, c, 4);
                b.power = pi * area * mean * (em.two_sided ? 2.0 : 1.0);
                b.bounds_power = pi * area * bound;
                b.two_sided = em.two_sided != 0;
                const D3 ns[4] = {dcross(dsub(p[1], p[0]), dsub(p[2], p[0])), dcross(dsub(p[1], p[0]), dsub(p[3], p[1])),
                                  dcross(dsub(p[3], p[2]), dsub(p[2], p[0])), dcross(dsub(p[3], p[2]), dsub(p[3], p[1]))};
                normal_cone(b, ns, 4);
                LightEntry L{};
                L.kind = 3;
                L.index = t;
                L.leaf = RP_NONE;
                L.rect = 1;
                L.inv_area = float(1.0 / area);
                const F3h ex = fsub(c[1], c[0]), ey = fsub(c[2], c[0]);
                const float pv[9] = {c[0].x, c[0].y, c[0].z, ex.x, ex.y, ex.z, ey.x, ey.y, ey.z};
                for (int k = 0; k < 3; ++k) {
                    L.p00[k] = pv[k];
                    L.ex[k] = pv[3 + k];
                    L.ey[k] = pv[6 + k];
                }
                // unit frame in binary32: u = e * (1 / |e|), z = u_x x u_y
                L.lx = std::sqrt(fdot(ex, ex));
                L.ly = std::sqrt(fdot(ey, ey));
                const float rx = 1.0f / L.lx, ry = 1.0f / L.ly;
                const F3h ux = {ex.x * rx, ex.y * rx, ex.z * rx}, uy = {ey.x * ry, ey.y * ry, ey.z * ry};
                const float uz[3] = {ux.y * uy.z - ux.z * uy.y, ux.z * uy.x - ux.x * uy.z, ux.x * uy.y - ux.y * uy.x};
                const float uxa[3] = {ux.x, ux.y, ux.z}, uya[3] = {uy.x, uy.y, uy.z};
                for (int k = 0; k < 3; ++k) {
                    L.ux[k] = uxa[k];
                    L.uy[k] = uya[k];
                    L.uz[k] = uz[k];
                }
                tri_light[t] = tri_light[t + 1] = uint32_t(out.lights.size());
                out.lights.push_back(L);
                in.push_back(b);
                ++t;
                continue;
            }
        }
        const F3h c[3] = {vert(ia[0]), vert(ia[1]), vert(ia[2])};
        const D3 n = dcross(dsub(dv(c[1]), dv(c[0])), dsub(dv(c[2]), dv(c[0])));
        const double area = 0.5 * dlen(n);
        if (!(area > 0)) continue;
        Bounds b;
        box_of(b, c, 3);
        b.power = pi * area * mean * (em.two_sided ? 2.0 : 1.0);
        b.bounds_power = pi * area * bound;
        b.two_sided = em.two_sided != 0;
        for (int a = 0; a < 3; ++a) b.dir[a] = float(n[a] / (2 * area));
        b.cos_outer = std::nextafter(1.0f, 0.0f);
        LightEntry L{};
        L.kind = 0;
        L.index = t;
        L.leaf = RP_NONE;
        L.inv_area = float(1.0 / area);
        tri_light[t] = uint32_t(out.lights.size());
        out.lights.push_back(L);
        in.push_back(b);
    }
    for (uint32_t i = 0; i < d.spot_count; ++i) {
        const rp_spot_light &sp = d.spots[i];
        Bounds b;
        for (int a = 0; a < 3; ++a) b.low[a] = b.high[a] = sp.position[a];
        b.power = 2 * pi * ((1 - double(sp.cos_falloff_start)) + (double(sp.cos_falloff_start) - sp.cos_falloff_end) / 2) *
                  spectrum_power(sp.intensity, d.spectra);
        b.bounds_power = 4 * pi * spectrum_max(sp.intensity, d.spectra);
        b.two_sided = false;
        float ce = float(std::cos(std::acos(double(sp.cos_falloff_end)) - std::acos(double(sp.cos_falloff_start))));
        if (ce == 1) ce = sp.cos_falloff_end != sp.cos_falloff_start ? 0.999f : std::nextafter(1.0f, 0.0f);
        for (int a = 0; a < 3; ++a) b.dir[a] = sp.direction[a];
        b.cos_outer = float(std::cos(std::min(pi, std::acos(std::clamp(double(sp.cos_falloff_start), -1.0, 1.0)) + 2e-6)));
        b.cos_outer = std::max(-1.0f, std::nextafter(b.cos_outer, -INFINITY));
        b.cos_emission = ce;
        LightEntry L{};
        L.kind = 1;
        L.index = i;
        L.leaf = RP_NONE;
        out.lights.push_back(L);
        in.push_back(b);
    }
    if (d.has_environment) {
        Bounds b;
        b.infinite = true;
        const double r = scene_radius;
        b.power = 4 * pi * pi * r * r * spectrum_power(d.environment, d.spectra);
        out.env_light = uint32_t(out.lights.size());
        LightEntry L{};
        L.kind = 2;
        L.leaf = RP_NONE;
        out.lights.push_back(L);
        in.push_back(b);
    }
    const size_t n = in.size();
    if (n == 0) return out;
    double total = 0, bounds_total = 0;
    std::vector<double> prob(n);
    std::vector<Bounds> bv = in;
    std::vector<uint32_t> finite;
    for (size_t i = 0; i < n; ++i) {
        bv[i].power = in[i].bounds_power < 0 ? in[i].power : in[i].bounds_power;
        if (bv[i].power == 0 && in[i].power > 0) bv[i].power = in[i].power;
        prob[i] = in[i].power == 0 && bv[i].power > 0 ? bv[i].power : in[i].power;
        total += prob[i];
        if (!in[i].infinite) bounds_total += bv[i].power;
        if (!in[i].infinite && bv[i].power > 0) finite.push_back(uint32_t(i));
    }
    // power alias table (Walker / Vose with small and large stacks)
    std::vector<uint32_t> small, large;
    for (size_t i = 0; i < n; ++i) {
        const double pr = total > 0 ? prob[i] / total : 1.0 / double(n);
        out.lights[i].pmf = float(pr);
        out.lights[i].threshold = 1;
        out.lights[i].alias = uint32_t(i);
        prob[i] = pr * double(n);
        (prob[i] < 1 ? small : large).push_back(uint32_t(i));
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t a = small.back();
        small.pop_back();
        const uint32_t b = large.back();
        large.pop_back();
        out.lights[a].threshold = float(prob[a]);
        out.lights[a].alias = b;
        prob[b] -= 1 - prob[a];
        (prob[b] < 1 ? small : large).push_back(b);
    }
    for (uint32_t i : small) out.lights[i].threshold = 1, out.lights[i].alias = i;
    for (uint32_t i : large) out.lights[i].threshold = 1, out.lights[i].alias = i;
    // light BVH: 12 centroid buckets per axis, splits 1..10
    auto bucket_of = [&](uint32_t id, int axis, const double lo[3], const double hi[3]) {
        const double c = (double(bv[id].low[axis]) + bv[id].high[axis]) / 2;
        return std::min(11, int(12 * (c - lo[axis]) / (hi[axis] - lo[axis])));
    };
    std::function<uint32_t(size_t, size_t, uint32_t, unsigned)> build = [&](size_t b, size_t e, uint32_t parent,
                                                                              unsigned depth) -> uint32_t {
        const uint32_t index = uint32_t(out.nodes.size());
        out.nodes.emplace_back();
        Bounds all;
        double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = b; k < e; ++k) {
            const Bounds &v = bv[finite[k]];
            all = unite(all, v);
            for (int a = 0; a < 3; ++a) {
                const double c = (double(v.low[a]) + v.high[a]) / 2;
                clo[a] = std::min(clo[a], c);
                chi[a] = std::max(chi[a], c);
            }
        }
        LightNodeHost node{};
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
            out.lights[finite[b]].leaf = index;
        } else {
            double best = INFINITY;
            int best_axis = 0, best_split = 0;
            bool found = false;
            if (depth < 48)
                for (int axis = 0; axis < 3; ++axis) {
                    if (clo[axis] == chi[axis]) continue;
                    Bounds bucket[12];
                    for (size_t k = b; k < e; ++k) {
                        const uint32_t id = finite[k];
                        const int q = bucket_of(id, axis, clo, chi);
                        bucket[q] = unite(bucket[q], bv[id]);
                    }
                    for (int split = 1; split < 11; ++split) {
                        Bounds left, right;
                        for (int k = 0; k <= split; ++k) left = unite(left, bucket[k]);
                        for (int k = split + 1; k < 12; ++k) right = unite(right, bucket[k]);
                        const double c = split_cost(left, all, axis) + split_cost(right, all, axis);
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
        out.nodes[index] = node;
        return index;
    };
    if (!finite.empty()) build(0, finite.size(), RP_NONE, 0);
    if (out.env_light != RP_NONE) {
        out.env_pmf = d.light_sampler == RP_LIGHT_SAMPLER_UNIFORM ? 1.0f / float(n)
                      : d.light_sampler == RP_LIGHT_SAMPLER_POWER  ? out.lights[out.env_light].pmf
                                                                    : 1.0f / float(1 + (out.nodes.empty() ? 0 : 1));
    }
    return out;
}

// ------------------------------------------------------------- scene copy
void HostScene::copy_from(const rp_scene_desc &d) {
    desc = d;
    positions.assign(d.positions, d.positions + 3 * size_t(d.vertex_count));
    indices.assign(d.indices, d.indices + 3 * size_t(d.triangle_count));
    tri_object.assign(d.tri_object, d.tri_object + d.triangle_count);
    objects.assign(d.objects, d.objects + d.object_count);
    materials.assign(d.materials, d.materials + d.material_count);
    emissions.assign(d.emissions, d.emissions + d.emission_count);
    media.assign(d.media, d.media + d.medium_count);
    spots.assign(d.spots, d.spots + d.spot_count);
    spectra.assign(d.spectra, d.spectra + d.spectra_count);
    desc.positions = positions.data();
    desc.indices = indices.data();
    desc.tri_object = tri_object.data();
    desc.objects = objects.data();
    desc.materials = materials.data();
    desc.emissions = emissions.data();
    desc.media = media.data();
    desc.spots = spots.data();
    desc.spectra = spectra.data();
}

static std::string check_spectrum(const rp_spectrum &s, const rp_scene_desc &d, const char *what) {
    if (s.kind > RP_SPECTRUM_KNOTS8) return std::string(what) + ": unknown spectrum kind";
    if (s.kind == RP_SPECTRUM_KNOTS8 && (uint64_t(s.table) + 8 > d.spectra_count))
        return std::string(what) + ": knot values outside the pool";
    if (s.kind == RP_SPECTRUM_TABLE && (uint64_t(s.table) + 471 > d.spectra_count))
        return std::string(what) + ": spectrum table outside the pool";
    if (!std::isfinite(s.scale)) return std::string(what) + ": non-finite scale";
    return {};
}

std::string validate_desc(const rp_scene_desc &d) {
    if (d.api_version != RP_API_VERSION) return "api_version mismatch";
    if (d.triangle_count && (!d.positions || !d.indices || !d.tri_object)) return "missing geometry arrays";
    if (d.object_count && !d.objects) return "missing objects";
    if (d.camera.width == 0 || d.camera.height == 0) return "empty film";
    if (uint64_t(d.camera.width) * d.camera.height > (1ull << 31)) return "film too large";
    if (std::max(d.camera.width, d.camera.height) > (1u << 25)) return "film extent beyond the Sobol table";
    // BSDF blocks 4.. and NEE blocks 133.. stay disjoint and inside 1024
    // dimensions (src/sampling_contract.h:16-28, dielectric_max_depth = 113).
    if (d.max_depth > 113) return "max_depth beyond the Sobol dimension layout (113)";
    if (d.light_sampler > RP_LIGHT_SAMPLER_UNIFORM) return "unknown light_sampler";
    if (d.transport > RP_TRANSPORT_SCALAR_DEMO) return "unknown transport";
    if (d.transport == RP_TRANSPORT_SCALAR_DEMO && !d.has_environment)
        return "the scalar demo model needs an environment (its only light)";
    for (uint32_t t = 0; t < d.triangle_count; ++t) {
        for (int k = 0; k < 3; ++k)
            if (d.indices[3 * t + k] >= d.vertex_count) return "triangle index out of range";
        if (d.tri_object[t] >= d.object_count) return "triangle object out of range";
    }
    for (uint32_t v = 0; v < 3 * d.vertex_count; ++v)
        if (!std::isfinite(d.positions[v])) return "non-finite vertex";
    for (uint32_t o = 0; o < d.object_count; ++o) {
        const rp_object &obj = d.objects[o];
        if (obj.material >= d.material_count) return "object material out of range";
        if (obj.emission != RP_NONE && obj.emission >= d.emission_count) return "object emission out of range";
        if (obj.medium_inside != RP_NONE && obj.medium_inside >= d.medium_count) return "medium out of range";
        if (obj.medium_outside != RP_NONE && obj.medium_outside >= d.medium_count) return "medium out of range";
    }
    for (uint32_t m = 0; m < d.material_count; ++m) {
        const rp_material &mt = d.materials[m];
        if (mt.kind > RP_MATERIAL_INTERFACE) return "unknown material kind";
        std::string e;
        if (!(e = check_spectrum(mt.reflectance, d, "reflectance")).empty()) return e;
        if (!(e = check_spectrum(mt.eta, d, "eta")).empty()) return e;
        if (!(e = check_spectrum(mt.k, d, "k")).empty()) return e;
        if (!(mt.roughness_u >= 0) || !(mt.roughness_v >= 0)) return "negative roughness";
    }
    for (uint32_t e = 0; e < d.emission_count; ++e) {
        std::string s = check_spectrum(d.emissions[e].radiance, d, "emission");
        if (!s.empty()) return s;
    }
    for (uint32_t m = 0; m < d.medium_count; ++m) {
        std::string s = check_spectrum(d.media[m].sigma_a, d, "sigma_a");
        if (s.empty()) s = check_spectrum(d.media[m].sigma_s, d, "sigma_s");
        if (!s.empty()) return s;
    }
    if (d.camera.medium != RP_NONE && d.camera.medium >= d.medium_count) return "camera medium out of range";
    if (d.has_environment) {
        std::string s = check_spectrum(d.environment, d, "environment");
        if (!s.empty()) return s;
    }
    // The scalar estimators scatter only at diffuse surfaces in vacuum
    // (shaders/scalar_transport.h:446-899); the demo model lights only by
    // the environment (docs/rendering.md:41-47).
    if (d.transport != RP_TRANSPORT_HERO) {
        for (uint32_t m = 0; m < d.material_count; ++m)
            if (d.materials[m].kind != RP_MATERIAL_DIFFUSE && d.materials[m].kind != RP_MATERIAL_INTERFACE)
                return "the scalar transport scatters only at diffuse surfaces";
        if (d.medium_count != 0 || d.camera.medium != RP_NONE) return "the scalar transport has no participating media";
    }
    if (d.transport == RP_TRANSPORT_SCALAR_DEMO) {
        if (d.spot_count != 0) return "the scalar demo model has no spot lights";
        for (uint32_t o = 0; o < d.object_count; ++o)
            if (d.objects[o].emission != RP_NONE) return "the scalar demo model has no emitters";
    }
    return {};
}

} // namespace rp
