// Seeded synthetic scene generators (include/rp_synthetic.h).
#include "rp_internal.hpp"

#include <cmath>
#include <map>
#include <new>

namespace rp {
namespace {

struct Rng { // splitmix64
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * 0x1p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

struct D3 {
    double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 normalize(D3 a) { return (1.0 / std::sqrt(dot(a, a))) * a; }

// Lattice value noise with quintic interpolation, fBm over octaves.
double lattice(int64_t x, int64_t y, uint64_t seed) {
    uint64_t h = uint64_t(x) * 0x8da6b343ull ^ uint64_t(y) * 0xd8163841ull ^ seed * 0xcb1ab31full;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return double(h >> 11) * 0x1p-53;
}
double value_noise(double x, double y, uint64_t seed) {
    double fx = std::floor(x), fy = std::floor(y);
    int64_t ix = int64_t(fx), iy = int64_t(fy);
    double tx = x - fx, ty = y - fy;
    auto q = [](double t) { return t * t * t * (t * (t * 6 - 15) + 10); };
    double sx = q(tx), sy = q(ty);
    double a = lattice(ix, iy, seed), b = lattice(ix + 1, iy, seed);
    double c = lattice(ix, iy + 1, seed), d = lattice(ix + 1, iy + 1, seed);
    return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sy;
}
double fbm(double x, double y, uint64_t seed, int octaves, double base) {
    double sum = 0, amp = 1, norm = 0, f = base;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * (2.0 * value_noise(x * f, y * f, seed + uint64_t(o) * 7919) - 1.0);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

struct Builder {
    std::vector<float> positions, spectra;
    std::vector<uint32_t> indices, tri_object;
    std::vector<rp_object> objects;
    std::vector<rp_material> materials;
    std::vector<rp_emission> emissions;
    std::vector<rp_medium> media;
    std::vector<rp_spot_light> spots;
    rp_scene_desc desc{};

    uint32_t vertex(D3 p) {
        positions.push_back(float(p.x));
        positions.push_back(float(p.y));
        positions.push_back(float(p.z));
        return uint32_t(positions.size() / 3 - 1);
    }
    void tri(uint32_t a, uint32_t b, uint32_t c, uint32_t obj) {
        indices.insert(indices.end(), {a, b, c});
        tri_object.push_back(obj);
    }
    // Quad with corners in order around the boundary; geometric normal
    // (cross(p0 - p2, p1 - p2)) of both triangles points along `normal`.
    void quad(D3 a, D3 b, D3 c, D3 d, D3 normal, uint32_t obj) {
        D3 n = cross(a - c, b - c);
        if (dot(n, normal) < 0) std::swap(b, d);
        uint32_t ia = vertex(a), ib = vertex(b), ic = vertex(c), id = vertex(d);
        tri(ia, ib, ic, obj);
        tri(ia, ic, id, obj);
    }
    uint32_t table(const std::array<float, 471> &t) {
        uint32_t off = uint32_t(spectra.size());
        spectra.insert(spectra.end(), t.begin(), t.end());
        return off;
    }
    uint32_t material(const rp_material &m) {
        materials.push_back(m);
        return uint32_t(materials.size() - 1);
    }
    uint32_t object(uint32_t mat, uint32_t emission = RP_NONE, uint32_t inside = RP_NONE, uint32_t outside = RP_NONE) {
        objects.push_back({mat, emission, inside, outside});
        return uint32_t(objects.size() - 1);
    }
    void finish() {
        desc.api_version = RP_API_VERSION;
        desc.vertex_count = uint32_t(positions.size() / 3);
        desc.triangle_count = uint32_t(tri_object.size());
        desc.object_count = uint32_t(objects.size());
        desc.positions = positions.data();
        desc.indices = indices.data();
        desc.tri_object = tri_object.data();
        desc.objects = objects.data();
        desc.material_count = uint32_t(materials.size());
        desc.emission_count = uint32_t(emissions.size());
        desc.medium_count = uint32_t(media.size());
        desc.spot_count = uint32_t(spots.size());
        desc.materials = materials.data();
        desc.emissions = emissions.data();
        desc.media = media.data();
        desc.spots = spots.data();
        desc.spectra = spectra.data();
        desc.spectra_count = uint32_t(spectra.size());
    }
};

rp_spectrum constant(float v) {
    rp_spectrum s{};
    s.kind = RP_SPECTRUM_CONSTANT;
    s.c[0] = v;
    s.scale = 1;
    return s;
}
rp_spectrum table_ref(uint32_t offset, float scale = 1.0f) {
    rp_spectrum s{};
    s.kind = RP_SPECTRUM_TABLE;
    s.table = offset;
    s.scale = scale;
    return s;
}
rp_material diffuse(rp_spectrum r) {
    rp_material m{};
    m.kind = RP_MATERIAL_DIFFUSE;
    m.reflectance = r;
    m.eta = constant(1);
    m.k = constant(0);
    return m;
}

// PBRT LookAt + perspective raster coefficients; fov spans the shorter axis.
rp_camera make_camera(D3 eye, D3 target, D3 up, double fov_deg, uint32_t w, uint32_t h) {
    D3 f = normalize(target - eye);
    D3 r = normalize(cross(f, up));
    D3 u = cross(r, f);
    double t = std::tan(fov_deg * 0.5 * 3.14159265358979323846 / 180.0);
    double sx = t, sy = t;
    if (w >= h) sx = t * double(w) / double(h);
    else sy = t * double(h) / double(w);
    D3 corner = f - sx * r + sy * u;
    D3 dx = (2 * sx) * r;
    D3 dy = (-2 * sy) * u;
    rp_camera c{};
    c.origin[0] = float(eye.x), c.origin[1] = float(eye.y), c.origin[2] = float(eye.z);
    c.corner[0] = float(corner.x), c.corner[1] = float(corner.y), c.corner[2] = float(corner.z);
    c.dx[0] = float(dx.x), c.dx[1] = float(dx.y), c.dx[2] = float(dx.z);
    c.dy[0] = float(dy.x), c.dy[1] = float(dy.y), c.dy[2] = float(dy.z);
    c.width = w;
    c.height = h;
    c.medium = RP_NONE;
    return c;
}

void box(Builder &b, D3 center_base, D3 size, double angle_deg, uint32_t obj) {
    double a = angle_deg * 3.14159265358979323846 / 180.0, ca = std::cos(a), sa = std::sin(a);
    auto P = [&](double x, double y, double z) {
        double lx = (x - 0.5) * size.x, lz = (z - 0.5) * size.z;
        return D3{center_base.x + ca * lx + sa * lz, center_base.y + y * size.y, center_base.z - sa * lx + ca * lz};
    };
    auto N = [&](double x, double y, double z) { return D3{ca * x + sa * z, y, -sa * x + ca * z}; };
    b.quad(P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1), N(0, 1, 0), obj);  // top
    b.quad(P(0, 0, 0), P(1, 0, 0), P(1, 0, 1), P(0, 0, 1), N(0, -1, 0), obj); // bottom
    b.quad(P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0), N(0, 0, -1), obj);
    b.quad(P(0, 0, 1), P(1, 0, 1), P(1, 1, 1), P(0, 1, 1), N(0, 0, 1), obj);
    b.quad(P(0, 0, 0), P(0, 0, 1), P(0, 1, 1), P(0, 1, 0), N(-1, 0, 0), obj);
    b.quad(P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0), N(1, 0, 0), obj);
}

void icosphere(Builder &b, D3 c, double r, int subdiv, uint32_t obj, uint64_t disp_seed = 0, double disp = 0) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto &p : v) p = normalize(p);
    std::vector<std::array<uint32_t, 3>> f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                               {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                               {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                               {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int s = 0; s < subdiv; ++s) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto m = [&](uint32_t a, uint32_t bb) {
            auto key = std::make_pair(std::min(a, bb), std::max(a, bb));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            v.push_back(normalize(0.5 * (v[a] + v[bb])));
            uint32_t id = uint32_t(v.size() - 1);
            mid[key] = id;
            return id;
        };
        std::vector<std::array<uint32_t, 3>> nf;
        nf.reserve(f.size() * 4);
        for (auto &tr : f) {
            uint32_t a = m(tr[0], tr[1]), bb = m(tr[1], tr[2]), cc = m(tr[2], tr[0]);
            nf.push_back({tr[0], a, cc});
            nf.push_back({tr[1], bb, a});
            nf.push_back({tr[2], cc, bb});
            nf.push_back({a, bb, cc});
        }
        f.swap(nf);
    }
    uint32_t base = uint32_t(b.positions.size() / 3);
    for (auto &p : v) {
        double rr = r;
        if (disp != 0) rr *= 1.0 + disp * fbm(std::atan2(p.z, p.x) * 2.0, std::acos(std::max(-1.0, std::min(1.0, p.y))) * 2.0, disp_seed, 6, 2.0);
        b.vertex(c + rr * p);
    }
    // Outward normals: cross(p0 - p2, p1 - p2) must point away from the centre.
    for (auto &tr : f) {
        D3 p0 = v[tr[0]], p1 = v[tr[1]], p2 = v[tr[2]];
        D3 n = cross(p0 - p2, p1 - p2);
        if (dot(n, p0 + p1 + p2) < 0) b.tri(base + tr[0], base + tr[2], base + tr[1], obj);
        else b.tri(base + tr[0], base + tr[1], base + tr[2], obj);
    }
}

void cornell_geometry(Builder &b, uint32_t white, uint32_t red, uint32_t green, uint32_t short_box,
                      uint32_t tall_box, uint32_t light) {
    const double S = 555;
    // floor, ceiling, back (white); left x = 555 red; right x = 0 green
    b.quad({0, 0, 0}, {S, 0, 0}, {S, 0, S}, {0, 0, S}, {0, 1, 0}, white);
    b.quad({0, S, 0}, {S, S, 0}, {S, S, S}, {0, S, S}, {0, -1, 0}, white);
    b.quad({0, 0, S}, {S, 0, S}, {S, S, S}, {0, S, S}, {0, 0, -1}, white);
    b.quad({S, 0, 0}, {S, 0, S}, {S, S, S}, {S, S, 0}, {-1, 0, 0}, red);
    b.quad({0, 0, 0}, {0, 0, S}, {0, S, S}, {0, S, 0}, {1, 0, 0}, green);
    box(b, {185.5, 0, 169}, {165, 165, 165}, -18, short_box);
    box(b, {368.5, 0, 351.5}, {165, 330, 165}, 15, tall_box);
    // 130 x 105 ceiling emitter facing down
    b.quad({213, 554, 227}, {343, 554, 227}, {343, 554, 332}, {213, 554, 332}, {0, -1, 0}, light);
}

void make_cornell(Builder &b, const rp_synthetic_options &o) {
    uint32_t w = o.width ? o.width : 256, h = o.height ? o.height : 256;
    uint32_t mw = b.material(diffuse(constant(0.73f)));
    uint32_t mr = b.material(diffuse(spectrum_from_rgb({0.65f, 0.05f, 0.05f})));
    uint32_t mg = b.material(diffuse(spectrum_from_rgb({0.12f, 0.45f, 0.15f})));
    rp_emission em{};
    em.radiance = table_ref(b.table(blackbody_table(6500, 1.0)), 17.0f);
    b.emissions.push_back(em);
    uint32_t ow = b.object(mw), orr = b.object(mr), og = b.object(mg);
    uint32_t os = b.object(mw), ot = b.object(mw), ol = b.object(mw, 0);
    cornell_geometry(b, ow, orr, og, os, ot, ol);
    b.desc.camera = make_camera({278, 273, -800}, {278, 273, 0}, {0, 1, 0}, 40, w, h);
    b.desc.max_depth = o.max_depth ? o.max_depth : 5;
    b.desc.hero_wavelengths = 0;
    b.desc.has_environment = 0;
}

// Per-object random material: 40% diffuse, 30% GGX conductor, 30% rough dielectric.
rp_material random_material(Rng &rng) {
    double u = rng.uniform();
    if (u < 0.4) {
        std::array<float, 3> rgb = {float(rng.uniform(0.1, 0.9)), float(rng.uniform(0.1, 0.9)), float(rng.uniform(0.1, 0.9))};
        return diffuse(spectrum_from_rgb(rgb));
    }
    rp_material m{};
    m.remap_roughness = 1;
    if (u < 0.7) {
        m.kind = RP_MATERIAL_CONDUCTOR;
        float r = float(rng.uniform(0.05, 0.5));
        m.roughness_u = m.roughness_v = r;
        m.eta.kind = RP_SPECTRUM_LINEAR;
        m.eta.c[0] = float(rng.uniform(0.2, 2.0));
        m.eta.c[1] = float(rng.uniform(0.2, 2.0));
        m.eta.scale = 1;
        m.k.kind = RP_SPECTRUM_LINEAR;
        m.k.c[0] = float(rng.uniform(1.0, 5.0));
        m.k.c[1] = float(rng.uniform(1.0, 5.0));
        m.k.scale = 1;
        m.reflectance = constant(0);
        return m;
    }
    m.kind = RP_MATERIAL_DIELECTRIC;
    float r = float(rng.uniform(0.0, 0.3));
    m.roughness_u = m.roughness_v = r;
    // Cauchy: eta(l) = A + B / l_um^2 with eta(550 nm) = 1.5, B = 0.0042 um^2.
    m.eta.kind = RP_SPECTRUM_CAUCHY;
    m.eta.c[0] = float(1.5 - 0.0042 / (0.55 * 0.55));
    m.eta.c[1] = 0.0042f;
    m.eta.scale = 1;
    m.k = constant(0);
    m.reflectance = constant(0);
    return m;
}

void make_terrain(Builder &b, uint64_t seed, const rp_synthetic_options &o, uint32_t grid, uint32_t spheres,
                  int subdiv, uint32_t w, uint32_t h) {
    Rng rng(seed);
    const double extent = 2.0, amp = 0.1 * extent;
    uint64_t nseed = rng.next();
    auto height = [&](double x, double z) { return amp * fbm(x, z, nseed, 6, 2.0); };
    // terrain object
    uint32_t mt = b.material(random_material(rng));
    uint32_t ot = b.object(mt);
    uint32_t base = uint32_t(b.positions.size() / 3);
    for (uint32_t j = 0; j < grid; ++j)
        for (uint32_t i = 0; i < grid; ++i) {
            double x = -1.0 + extent * double(i) / double(grid - 1);
            double z = -1.0 + extent * double(j) / double(grid - 1);
            b.vertex({x, height(x, z), z});
        }
    for (uint32_t j = 0; j + 1 < grid; ++j)
        for (uint32_t i = 0; i + 1 < grid; ++i) {
            uint32_t v00 = base + j * grid + i, v10 = v00 + 1, v01 = v00 + grid, v11 = v01 + 1;
            // upward normals: cross(p0 - p2, p1 - p2) has +y
            b.tri(v00, v01, v11, ot);
            b.tri(v00, v11, v10, ot);
        }
    for (uint32_t s = 0; s < spheres; ++s) {
        double r = rng.uniform(0.02, 0.1);
        double x = rng.uniform(-0.85, 0.85), z = rng.uniform(-0.85, 0.85);
        uint32_t ms = b.material(random_material(rng));
        uint32_t os = b.object(ms);
        icosphere(b, {x, height(x, z) + 0.5 * r, z}, r, subdiv, os);
    }
    // 16 blackbody quad emitters facing down
    uint32_t ml = b.material(diffuse(constant(0.5f)));
    for (int q = 0; q < 16; ++q) {
        double x = rng.uniform(-0.8, 0.8), z = rng.uniform(-0.8, 0.8);
        double y = rng.uniform(0.45, 0.7), half = 0.5 * rng.uniform(0.05, 0.15);
        double kelvin = rng.uniform(3000, 8000);
        rp_emission em{};
        em.radiance = table_ref(b.table(blackbody_table(kelvin, 1.0)), 12.0f);
        b.emissions.push_back(em);
        uint32_t ol = b.object(ml, uint32_t(b.emissions.size() - 1));
        b.quad({x - half, y, z - half}, {x + half, y, z - half}, {x + half, y, z + half}, {x - half, y, z + half},
               {0, -1, 0}, ol);
    }
    b.desc.has_environment = 1;
    b.desc.environment = constant(0.5f);
    b.desc.camera = make_camera({0, 0.85, -1.9}, {0, 0, 0.05}, {0, 1, 0}, 45, w, h);
    b.desc.max_depth = o.max_depth ? o.max_depth : 8;
    b.desc.hero_wavelengths = 0;
}

std::array<float, 471> binned_table(const std::vector<float> &bins) {
    std::array<float, 471> t{};
    uint32_t n = uint32_t(bins.size());
    for (uint32_t i = 0; i < 471; ++i) {
        uint32_t bin = std::min(n - 1, uint32_t((double(i) * n) / 471.0));
        t[i] = bins[bin];
    }
    return t;
}

void make_volume(Builder &b, uint64_t seed, const rp_synthetic_options &o) {
    Rng rng(seed);
    uint32_t w = o.width ? o.width : 512, h = o.height ? o.height : 512;
    int subdiv = o.volume_subdivision ? int(o.volume_subdivision) : 6;
    std::vector<float> sa(8), ss(8);
    for (int i = 0; i < 8; ++i) sa[i] = float(rng.uniform(0.05, 2.0));
    for (int i = 0; i < 8; ++i) ss[i] = float(rng.uniform(0.05, 2.0));
    rp_medium med{};
    med.sigma_a = table_ref(b.table(binned_table(sa)));
    med.sigma_s = table_ref(b.table(binned_table(ss)));
    med.g = o.hg_g;
    b.media.push_back(med);
    rp_material glass{};
    glass.kind = RP_MATERIAL_DIELECTRIC;
    glass.eta = constant(1.33f);
    glass.k = constant(0);
    glass.reflectance = constant(0);
    glass.remap_roughness = 1;
    uint32_t mg = b.material(glass);
    uint32_t os = b.object(mg, RP_NONE, 0, RP_NONE);
    icosphere(b, {0, 0, 0}, 1.0, subdiv, os, rng.next(), 0.1);
    rp_spot_light sp{};
    D3 pos = {2.5, 3.0, -2.5}, dir = normalize(D3{0, 0, 0} - pos);
    sp.position[0] = float(pos.x), sp.position[1] = float(pos.y), sp.position[2] = float(pos.z);
    sp.direction[0] = float(dir.x), sp.direction[1] = float(dir.y), sp.direction[2] = float(dir.z);
    sp.cos_falloff_end = float(std::cos(30.0 * 3.14159265358979323846 / 180.0));
    sp.cos_falloff_start = float(std::cos(25.0 * 3.14159265358979323846 / 180.0));
    sp.intensity = table_ref(b.table(blackbody_table(4000, 1.0)), 40.0f);
    b.spots.push_back(sp);
    b.desc.has_environment = 1;
    b.desc.environment = constant(0.2f);
    b.desc.camera = make_camera({0, 0.6, -3.5}, {0, 0, 0}, {0, 1, 0}, 40, w, h);
    b.desc.max_depth = o.max_depth ? o.max_depth : 32;
    b.desc.hero_wavelengths = 0;
}

void make_inverse(Builder &b, uint64_t seed, const rp_synthetic_options &o) {
    Rng rng(seed);
    uint32_t w = o.width ? o.width : 256, h = o.height ? o.height : 256;
    uint32_t bins = o.bins ? o.bins : 16;
    auto wall = [&](std::array<float, 3> rgb) {
        rp_spectrum s = spectrum_from_rgb(rgb);
        std::vector<float> v(bins);
        for (uint32_t i = 0; i < bins; ++i) {
            double acc = 0;
            int n = 0;
            for (uint32_t l = 0; l < 471; ++l)
                if (std::min(bins - 1, uint32_t((double(l) * bins) / 471.0)) == i) {
                    acc += spectrum_eval_host(s, float(360 + l), nullptr);
                    ++n;
                }
            v[i] = float(acc / std::max(n, 1));
        }
        return diffuse(table_ref(b.table(binned_table(v))));
    };
    uint32_t mw = b.material(wall({0.73f, 0.73f, 0.73f}));
    uint32_t mr = b.material(wall({0.65f, 0.05f, 0.05f}));
    uint32_t mgr = b.material(wall({0.12f, 0.45f, 0.15f}));
    auto metal = [&](float rough) {
        rp_material m{};
        m.kind = RP_MATERIAL_CONDUCTOR;
        m.remap_roughness = 1;
        m.roughness_u = m.roughness_v = rough;
        m.eta.kind = RP_SPECTRUM_LINEAR;
        m.eta.c[0] = 1.3f;
        m.eta.c[1] = 0.2f;
        m.eta.scale = 1;
        m.k.kind = RP_SPECTRUM_LINEAR;
        m.k.c[0] = 1.9f;
        m.k.c[1] = 5.0f;
        m.k.scale = 1;
        m.reflectance = constant(0);
        return m;
    };
    uint32_t ms = b.material(metal(0.2f)), mt = b.material(metal(0.35f));
    uint32_t ml = b.material(diffuse(constant(0.73f)));
    rp_material fog{};
    fog.kind = RP_MATERIAL_INTERFACE;
    fog.reflectance = constant(0);
    fog.eta = constant(1);
    fog.k = constant(0);
    uint32_t mf = b.material(fog);
    rp_emission em{};
    em.radiance = table_ref(b.table(blackbody_table(6500, 1.0)), 17.0f);
    b.emissions.push_back(em);
    std::vector<float> ssb(8);
    for (int i = 0; i < 8; ++i) ssb[i] = float(rng.uniform(0.002, 0.006));
    rp_medium med{};
    med.sigma_a = constant(0.0005f);
    med.sigma_s = table_ref(b.table(binned_table(ssb)));
    med.g = 0.2f;
    b.media.push_back(med);
    uint32_t ow = b.object(mw), orr = b.object(mr), og = b.object(mgr);
    uint32_t os = b.object(ms), ot = b.object(mt), ol = b.object(ml, 0);
    cornell_geometry(b, ow, orr, og, os, ot, ol);
    uint32_t of = b.object(mf, RP_NONE, 0, RP_NONE);
    box(b, {277.5, 360, 277.5}, {455, 140, 455}, 0, of);
    b.desc.camera = make_camera({278, 273, -800}, {278, 273, 0}, {0, 1, 0}, 40, w, h);
    b.desc.max_depth = o.max_depth ? o.max_depth : 5;
    b.desc.hero_wavelengths = 0;
}

// The reference's demo scene (src/render.cpp:2060-2106): floor and two
// walls as quads triangulated (0,1,2),(0,2,3) in the given winding, one unit
// cube mesh instanced three times by box_transform (src/render.cpp:821-827,
// rows {c w, 0, s d, x}, {0, h, 0, y}, {-s w, 0, c d, z}), six diffuse
// materials as reflectance at 360/420/470/520/570/620/700/830 nm, a unit
// environment and a 42-degree vertical field of view from (4.1, 3, 6.2).
void make_demo(Builder &b, const rp_synthetic_options &o) {
    const uint32_t w = o.width ? o.width : 256, h = o.height ? o.height : 192;
    static const float knots[6][8] = {
        {0.54f, 0.54f, 0.55f, 0.58f, 0.61f, 0.63f, 0.65f, 0.65f},
        {0.69f, 0.68f, 0.67f, 0.65f, 0.64f, 0.62f, 0.61f, 0.61f},
        {0.34f, 0.34f, 0.36f, 0.43f, 0.55f, 0.61f, 0.63f, 0.63f},
        {0.10f, 0.10f, 0.16f, 0.52f, 0.40f, 0.09f, 0.06f, 0.06f},
        {0.07f, 0.08f, 0.09f, 0.13f, 0.39f, 0.73f, 0.72f, 0.72f},
        {0.32f, 0.47f, 0.72f, 0.23f, 0.11f, 0.08f, 0.09f, 0.09f},
    };
    uint32_t obj[6];
    for (int m = 0; m < 6; ++m) {
        rp_spectrum r{};
        r.kind = RP_SPECTRUM_KNOTS8;
        r.table = uint32_t(b.spectra.size());
        r.scale = 1;
        b.spectra.insert(b.spectra.end(), knots[m], knots[m] + 8);
        obj[m] = b.object(b.material(diffuse(r)));
    }
    using P3 = std::array<float, 3>;
    auto mesh_quad = [&](const P3 (&c)[4], uint32_t object) {
        uint32_t v[4];
        for (int i = 0; i < 4; ++i) {
            b.positions.insert(b.positions.end(), c[i].begin(), c[i].end());
            v[i] = uint32_t(b.positions.size() / 3 - 1);
        }
        b.tri(v[0], v[1], v[2], object);
        b.tri(v[0], v[2], v[3], object);
    };
    mesh_quad({{-3, 0, -3}, {-3, 0, 3}, {3, 0, 3}, {3, 0, -3}}, obj[0]);
    mesh_quad({{-3, 0, -3}, {3, 0, -3}, {3, 3.2f, -3}, {-3, 3.2f, -3}}, obj[1]);
    mesh_quad({{-3, 0, 3}, {-3, 0, -3}, {-3, 3.2f, -3}, {-3, 3.2f, 3}}, obj[2]);
    static const float cube[8][3] = {{-0.5f, 0, -0.5f}, {0.5f, 0, -0.5f}, {0.5f, 0, 0.5f}, {-0.5f, 0, 0.5f},
                                     {-0.5f, 1, -0.5f}, {0.5f, 1, -0.5f}, {0.5f, 1, 0.5f}, {-0.5f, 1, 0.5f}};
    static const uint32_t cube_tris[12][3] = {{0, 1, 2}, {0, 2, 3}, {4, 6, 5}, {4, 7, 6}, {0, 4, 5}, {0, 5, 1},
                                              {1, 5, 6}, {1, 6, 2}, {2, 6, 7}, {2, 7, 3}, {3, 7, 4}, {3, 4, 0}};
    struct DemoBox {
        float offset[3], scale[3], angle;
    };
    static const DemoBox boxes[3] = {{{-1.25f, -0.05f, -0.6f}, {1.25f, 2.25f, 1.25f}, 0.18f},
                                     {{0.25f, -0.05f, 0.8f}, {1.45f, 1.08f, 1.3f}, -0.24f},
                                     {{1.4f, -0.05f, -1.5f}, {1.07f, 1.55f, 1.06f}, 0.44f}};
    for (int i = 0; i < 3; ++i) {
        const DemoBox &bx = boxes[i];
        const float c = std::cos(bx.angle), s = std::sin(bx.angle);
        const float m[12] = {c * bx.scale[0], 0, s * bx.scale[2], bx.offset[0], 0, bx.scale[1], 0, bx.offset[1],
                             -s * bx.scale[0], 0, c * bx.scale[2], bx.offset[2]};
        const uint32_t base = uint32_t(b.positions.size() / 3);
        for (const auto &p : cube)
            for (int r = 0; r < 3; ++r)
                b.positions.push_back(m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3]);
        for (const auto &t : cube_tris) b.tri(base + t[0], base + t[1], base + t[2], obj[3 + i]);
    }
    // Vertical field of view whatever the aspect (the demo camera).
    const D3 eye = {4.1f, 3.0f, 6.2f}, target = {0.0f, 1.0f, -0.45f}, up = {0, 1, 0};
    const D3 f = normalize(target - eye), r = normalize(cross(f, up)), u = cross(r, f);
    const double th = std::tan(42.0 * 0.5 * 3.14159265358979323846 / 180.0), asp = double(w) / double(h);
    const D3 corner = f - (th * asp) * r + th * u, dx = (2 * th * asp) * r, dy = (-2 * th) * u;
    rp_camera cam{};
    cam.origin[0] = float(eye.x), cam.origin[1] = float(eye.y), cam.origin[2] = float(eye.z);
    cam.corner[0] = float(corner.x), cam.corner[1] = float(corner.y), cam.corner[2] = float(corner.z);
    cam.dx[0] = float(dx.x), cam.dx[1] = float(dx.y), cam.dx[2] = float(dx.z);
    cam.dy[0] = float(dy.x), cam.dy[1] = float(dy.y), cam.dy[2] = float(dy.z);
    cam.width = w;
    cam.height = h;
    cam.medium = RP_NONE;
    b.desc.camera = cam;
    b.desc.has_environment = 1;
    b.desc.environment = constant(1.0f);
    b.desc.transport = RP_TRANSPORT_SCALAR_DEMO;
    b.desc.max_depth = o.max_depth ? o.max_depth : 113;
    b.desc.hero_wavelengths = 0;
}

} // namespace
} // namespace rp

struct rp_synthetic {
    rp::Builder b;
};

extern "C" {

int rp_synthetic_create(uint32_t kind, uint64_t seed, const rp_synthetic_options *options, rp_synthetic **out) {
    using namespace rp;
    if (!out) {
        set_error("rp_synthetic_create: null output");
        return RP_INVALID;
    }
    *out = nullptr;
    rp_synthetic_options o{};
    if (options) o = *options;
    try {
        auto *s = new rp_synthetic();
        Builder &b = s->b;
        switch (kind) {
        case RP_SYNTH_CORNELL: make_cornell(b, o); break;
        case RP_SYNTH_TERRAIN: {
            uint32_t grid = o.grid ? o.grid : 1024;
            uint32_t spheres = o.spheres ? o.spheres : 64;
            int sub = o.sphere_subdivision ? int(o.sphere_subdivision) : 4;
            make_terrain(b, seed, o, grid, spheres, sub, o.width ? o.width : 1920, o.height ? o.height : 1080);
            break;
        }
        case RP_SYNTH_BATCH: {
            // Triangle budget: ~80% heightfield, ~20% subdivision-3 icospheres (1280 tris each).
            double target = o.target_triangles ? double(o.target_triangles) : 50000.0;
            uint32_t spheres = std::max(1u, uint32_t(std::lround(0.2 * target / 1280.0)));
            double hf = std::max(200.0, target - 1280.0 * spheres);
            uint32_t grid = std::max(4u, uint32_t(std::lround(std::sqrt(hf / 2.0))) + 1);
            make_terrain(b, seed, o, grid, spheres, 3, o.width ? o.width : 256, o.height ? o.height : 256);
            break;
        }
        case RP_SYNTH_VOLUME: make_volume(b, seed, o); break;
        case RP_SYNTH_INVERSE: make_inverse(b, seed, o); break;
        case RP_SYNTH_DEMO: make_demo(b, o); break;
        default:
            delete s;
            set_error("rp_synthetic_create: unknown kind");
            return RP_INVALID;
        }
        b.finish();
        *out = s;
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        set_error(std::string("rp_synthetic_create: ") + e.what());
        return RP_FAILURE;
    }
}

const rp_scene_desc *rp_synthetic_desc(const rp_synthetic *s) { return s ? &s->b.desc : nullptr; }
void rp_synthetic_destroy(rp_synthetic *s) { delete s; }

} // extern "C"
