// Host services of the rypath_amd library: errors, data tables, spectra,
// light tables and scene validation.
#include "rp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
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

static float avg_spectrum(const rp_spectrum &s, const float *pool) {
    double acc = 0;
    for (int i = 0; i < 471; ++i) acc += double(spectrum_eval_host(s, float(360 + i), pool));
    return float(acc / 471.0);
}

std::vector<LightEntry> build_light_table(const rp_scene_desc &d, float scene_radius, std::vector<uint32_t> &tri_light) {
    const double pi = double(3.14159265358979323846f);
    std::vector<LightEntry> lights;
    std::vector<double> power;
    tri_light.assign(d.triangle_count, RP_NONE);
    for (uint32_t t = 0; t < d.triangle_count; ++t) {
        const rp_object &obj = d.objects[d.tri_object[t]];
        if (obj.emission == RP_NONE) continue;
        const rp_emission &em = d.emissions[obj.emission];
        const float *a = d.positions + 3 * size_t(d.indices[3 * t]);
        const float *b = d.positions + 3 * size_t(d.indices[3 * t + 1]);
        const float *c = d.positions + 3 * size_t(d.indices[3 * t + 2]);
        double ex = double(b[0]) - a[0], ey = double(b[1]) - a[1], ez = double(b[2]) - a[2];
        double fx = double(c[0]) - a[0], fy = double(c[1]) - a[1], fz = double(c[2]) - a[2];
        double cx = ey * fz - ez * fy, cy = ez * fx - ex * fz, cz = ex * fy - ey * fx;
        double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        if (!(area > 0)) continue;
        double p = pi * area * double(avg_spectrum(em.radiance, d.spectra)) * (em.two_sided ? 2.0 : 1.0);
        tri_light[t] = uint32_t(lights.size());
        lights.push_back({0u, t, 0.0f, 0.0f, float(1.0 / area), {0, 0, 0}});
        power.push_back(p);
    }
    for (uint32_t i = 0; i < d.spot_count; ++i) {
        const rp_spot_light &sp = d.spots[i];
        double p = double(avg_spectrum(sp.intensity, d.spectra)) * 2.0 * pi *
                   (1.0 - 0.5 * (double(sp.cos_falloff_start) + double(sp.cos_falloff_end)));
        lights.push_back({1u, i, 0.0f, 0.0f, 0.0f, {0, 0, 0}});
        power.push_back(p);
    }
    if (d.has_environment) {
        double r = scene_radius;
        double p = pi * r * r * 4.0 * pi * double(avg_spectrum(d.environment, d.spectra));
        lights.push_back({2u, 0u, 0.0f, 0.0f, 0.0f, {0, 0, 0}});
        power.push_back(p);
    }
    double total = 0;
    for (double p : power) total += p;
    double cum = 0;
    for (size_t i = 0; i < lights.size(); ++i) {
        cum += power[i];
        lights[i].pmf = total > 0 ? float(power[i] / total) : 0.0f;
        lights[i].cdf = total > 0 ? float(cum / total) : 0.0f;
    }
    if (!lights.empty() && total > 0) lights.back().cdf = 1.0f;
    return lights;
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
    if (s.kind > RP_SPECTRUM_LINEAR) return std::string(what) + ": unknown spectrum kind";
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
    return {};
}

} // namespace rp
