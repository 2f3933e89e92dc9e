// PBRT-v4 scene-file subset -> rp_scene_desc (include/rp_io.h).
//
// Counterpart of the reference's PBRT front end (src/pbrt.hpp:105-241 check,
// src/pbrt_render.hpp translate) restricted to what the renderer supports:
//   Camera "perspective", Film (x/yresolution, filename), Sampler "sobol"
//   (pixelsamples, seed, randomization), Integrator "path"/"volpath"
//   (maxdepth), LookAt / Translate / Scale / Rotate / Transform /
//   ConcatTransform / Identity, AttributeBegin/End, Material /
//   MakeNamedMaterial / NamedMaterial ("diffuse", "conductor", "dielectric",
//   "interface"), Shape "trianglemesh" and "plymesh" (ASCII or binary
//   little-endian PLY), AreaLightSource "diffuse", LightSource "infinite"
//   (constant) and "spot", MakeNamedMedium "homogeneous", MediumInterface,
//   ReverseOrientation, Include.  Anything else is an error naming the
//   directive.
#include "rp_internal.hpp"
#include "../../include/rp_io.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace rp {
namespace {

using M4 = std::array<double, 16>; // row-major

M4 identity() { return {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }
M4 mul(const M4 &a, const M4 &b) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = s;
        }
    return r;
}
bool invert(const M4 &m, M4 &out) {
    // Gauss-Jordan with partial pivoting
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (std::fabs(a[piv][c]) < 1e-300) return false;
        for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[piv][j]);
        double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[i * 4 + j] = a[i][j + 4];
    return true;
}
std::array<double, 3> xform_point(const M4 &m, double x, double y, double z) {
    double w = m[12] * x + m[13] * y + m[14] * z + m[15];
    return {(m[0] * x + m[1] * y + m[2] * z + m[3]) / w, (m[4] * x + m[5] * y + m[6] * z + m[7]) / w,
            (m[8] * x + m[9] * y + m[10] * z + m[11]) / w};
}
std::array<double, 3> xform_vector(const M4 &m, double x, double y, double z) {
    return {m[0] * x + m[1] * y + m[2] * z, m[4] * x + m[5] * y + m[6] * z, m[8] * x + m[9] * y + m[10] * z};
}

// ------------------------------------------------------------- tokens
struct Token {
    enum Kind { word, string, number, open, close } kind;
    std::string text;
    double value = 0;
};

struct Lexer {
    std::string src, file;
    size_t pos = 0;
    int line = 1;
    int token_line = 1;     // line of the last token read
    int directive_line = 0; // line of the directive being translated (errors)
    bool next(Token &t) {
        for (;;) {
            while (pos < src.size() && std::isspace(static_cast<unsigned char>(src[pos]))) {
                if (src[pos] == '\n') ++line;
                ++pos;
            }
            if (pos < src.size() && src[pos] == '#') {
                while (pos < src.size() && src[pos] != '\n') ++pos;
                continue;
            }
            break;
        }
        if (pos >= src.size()) return false;
        token_line = line;
        char c = src[pos];
        if (c == '[') {
            ++pos;
            t = {Token::open, "[", 0};
            return true;
        }
        if (c == ']') {
            ++pos;
            t = {Token::close, "]", 0};
            return true;
        }
        if (c == '"') {
            size_t e = src.find('"', pos + 1);
            if (e == std::string::npos) fail("unterminated string");
            t = {Token::string, src.substr(pos + 1, e - pos - 1), 0};
            pos = e + 1;
            return true;
        }
        size_t b = pos;
        while (pos < src.size() && !std::isspace(static_cast<unsigned char>(src[pos])) && src[pos] != '[' &&
               src[pos] != ']' && src[pos] != '"' && src[pos] != '#')
            ++pos;
        std::string w = src.substr(b, pos - b);
        char *end = nullptr;
        double v = std::strtod(w.c_str(), &end);
        if (end && *end == 0 && !w.empty() && (std::isdigit(static_cast<unsigned char>(w[0])) || w[0] == '-' ||
                                               w[0] == '+' || w[0] == '.')) {
            t = {Token::number, w, v};
        } else {
            t = {Token::word, w, 0};
        }
        return true;
    }
    [[noreturn]] void fail(const std::string &m) const {
        throw std::runtime_error(file + ":" + std::to_string(directive_line ? directive_line : token_line) + ": " + m);
    }
};

struct Param {
    std::string type, name;
    std::vector<double> nums;
    std::vector<std::string> strs;
    bool used = false;
};
struct ParamSet {
    std::vector<Param> p;
    Param *find(const std::string &name) {
        for (auto &x : p)
            if (x.name == name) {
                x.used = true;
                return &x;
            }
        return nullptr;
    }
    double f(const std::string &n, double def) {
        Param *x = find(n);
        return x && !x->nums.empty() ? x->nums[0] : def;
    }
    std::string s(const std::string &n, const std::string &def) {
        Param *x = find(n);
        return x && !x->strs.empty() ? x->strs[0] : def;
    }
    bool b(const std::string &n, bool def) {
        Param *x = find(n);
        if (!x) return def;
        if (!x->strs.empty()) return x->strs[0] == "true";
        return !x->nums.empty() ? x->nums[0] != 0 : def;
    }
};

struct Graphics {
    M4 ctm = identity();
    std::string material = "";
    bool reverse = false;
    bool area = false;
    rp_emission emission{};
    std::string inside, outside;
};

// ------------------------------------------------------------- translator
struct Translator {
    std::vector<float> positions, spectra;
    std::vector<uint32_t> indices, tri_object;
    std::vector<rp_object> objects;
    std::vector<rp_material> materials;
    std::vector<rp_emission> emissions;
    std::vector<rp_medium> media;
    std::vector<rp_spot_light> spots;
    rp_scene_desc desc{};
    rp_pbrt_info info{};
    uint32_t light_sampler = RP_LIGHT_SAMPLER_BVH;
    std::map<std::string, uint32_t> named_materials, named_media;
    std::vector<Graphics> stack;
    Graphics g;
    M4 camera_world = identity();
    double fov = 90;
    bool have_camera = false, in_world = false;
    std::string base_dir;
    int include_depth = 0;

    Translator() {
        info.width = 1280;
        info.height = 720;
        info.spp = 16;
        info.randomization = RP_RANDOMIZE_FASTOWEN;
        info.seed = 0;
        info.max_depth = 5;
        rp_material d{};
        d.kind = RP_MATERIAL_DIFFUSE;
        d.reflectance = constant(0.5f);
        d.eta = constant(1.0f);
        d.k = constant(0.0f);
        materials.push_back(d); // PBRT default material: diffuse 0.5
        g.material = "";
    }

    static rp_spectrum constant(float v) {
        rp_spectrum s{};
        s.kind = RP_SPECTRUM_CONSTANT;
        s.c[0] = v;
        s.scale = 1;
        return s;
    }
    uint32_t table(const std::array<float, 471> &t) {
        uint32_t off = uint32_t(spectra.size());
        spectra.insert(spectra.end(), t.begin(), t.end());
        return off;
    }
    // Piecewise-linear knots (l0 v0 l1 v1 ..., increasing l) sampled at
    // 360..830 nm, zero outside [l0, l_last] (PBRT PiecewiseLinearSpectrum;
    // reference resampling src/render.cpp:2164-2183).
    static std::array<float, 471> knots_dense(const Param *p, const Lexer &lx) {
        const auto &v = p->nums;
        if (v.size() < 2 || v.size() % 2) lx.fail("spectrum '" + p->name + "' needs wavelength/value pairs");
        for (size_t k = 2; k < v.size(); k += 2)
            if (!(v[k] > v[k - 2])) lx.fail("spectrum '" + p->name + "' wavelengths must increase");
        std::array<float, 471> out{};
        const size_t n = v.size() / 2;
        for (int i = 0; i < 471; ++i) {
            const double l = 360.0 + i;
            if (l < v[0] || l > v[2 * (n - 1)]) continue;
            size_t k = 0;
            while (k + 1 < n && v[2 * (k + 1)] < l) ++k;
            if (k + 1 == n || v[2 * k] == l) {
                out[i] = float(v[2 * k + 1]);
                continue;
            }
            const double l0 = v[2 * k], l1 = v[2 * k + 2], a = v[2 * k + 1], b = v[2 * k + 3];
            const double left = l - l0, right = l1 - l, w = l1 - l0;
            // anchored at the nearer knot so tiny contributions survive
            out[i] = float(right < left ? b + (a - b) * (right / w) : a + (b - a) * (left / w));
        }
        return out;
    }
    static std::array<float, 3> rgb_of(const Param *p, const Lexer &lx) {
        if (p->nums.size() != 3) lx.fail("rgb '" + p->name + "' needs three values");
        std::array<float, 3> rgb = {float(p->nums[0]), float(p->nums[1]), float(p->nums[2])};
        for (float c : rgb)
            if (!(c >= 0) || !std::isfinite(c)) lx.fail("rgb '" + p->name + "' must be finite and nonnegative");
        return rgb;
    }
    // Reflectances (PBRT SpectrumType::Albedo): rgb in [0,1] -> sigmoid.
    rp_spectrum albedo(const Param *p, const Lexer &lx) {
        if (p->type == "float") return constant(float(p->nums.at(0)));
        if (p->type == "rgb") {
            auto rgb = rgb_of(p, lx);
            if (rgb[0] > 1 || rgb[1] > 1 || rgb[2] > 1) lx.fail("reflectance rgb above 1");
            return spectrum_from_rgb(rgb);
        }
        if (p->type == "spectrum" && p->strs.empty()) return table_ref(table(knots_dense(p, lx)));
        if (p->type == "blackbody") return table_ref(table(blackbody_table(p->nums.at(0), 1.0)));
        lx.fail("unsupported spectrum for '" + p->name + "'");
    }
    // Unbounded quantities (eta, k, sigma_a, sigma_s): rgb -> 2m * sigmoid(rgb / 2m).
    rp_spectrum unbounded(const Param *p, const Lexer &lx) {
        if (p->type == "float") return constant(float(p->nums.at(0)));
        if (p->type == "rgb") {
            auto rgb = rgb_of(p, lx);
            float m = std::max(rgb[0], std::max(rgb[1], rgb[2]));
            if (m == 0) return constant(0);
            float s = 2 * m;
            rp_spectrum out = spectrum_from_rgb({rgb[0] / s, rgb[1] / s, rgb[2] / s});
            out.scale = s;
            return out;
        }
        if (p->type == "spectrum" && p->strs.empty()) return table_ref(table(knots_dense(p, lx)));
        if (p->type == "blackbody") return table_ref(table(blackbody_table(p->nums.at(0), 1.0)));
        if (p->type == "spectrum") lx.fail("named spectrum '" + p->strs[0] + "' is not supported");
        lx.fail("unsupported spectrum for '" + p->name + "'");
    }
    // Emission (PBRT SpectrumType::Illuminant + photometric normalisation):
    // dense = D65 * sigmoid(rgb / 2m) * 2m for rgb, the knots, or a
    // peak-normalised blackbody; the result is dense * scale / <dense_base, Y>
    // where dense_base is D65 for rgb (src/pbrt_render.cpp:940-1033).
    rp_spectrum emission(const Param *p, double scale, const Lexer &lx) {
        std::array<float, 471> dense = d65();
        double norm = 0;
        if (!p || p->type == "rgb") {
            norm = y_inner_product(dense);
            if (p) {
                auto rgb = rgb_of(p, lx);
                float m = std::max(rgb[0], std::max(rgb[1], rgb[2]));
                float s = 2 * m;
                if (s > 0) {
                    rp_spectrum tint = spectrum_from_rgb({rgb[0] / s, rgb[1] / s, rgb[2] / s});
                    for (int i = 0; i < 471; ++i) dense[i] *= spectrum_eval_host(tint, 360.0f + float(i), nullptr);
                }
                scale *= s;
            }
        } else if (p->type == "blackbody") {
            dense = blackbody_table(p->nums.at(0), 1.0);
            norm = y_inner_product(dense);
        } else if (p->type == "spectrum" && p->strs.empty()) {
            dense = knots_dense(p, lx);
            norm = y_inner_product(dense);
        } else if (p->type == "float") {
            // a constant spectrum: photometric normalisation of the constant
            dense.fill(float(p->nums.at(0)));
            norm = y_inner_product(dense);
        } else {
            lx.fail("unsupported emission spectrum for '" + p->name + "'");
        }
        if (!(norm > 0) || !std::isfinite(norm)) lx.fail("emission spectrum needs a positive CIE Y inner product");
        for (auto &v : dense) v = float(double(v) * scale / norm);
        return table_ref(table(dense));
    }
    static rp_spectrum table_ref(uint32_t off) {
        rp_spectrum s{};
        s.kind = RP_SPECTRUM_TABLE;
        s.table = off;
        s.scale = 1;
        return s;
    }

    rp_material make_material(const std::string &type, ParamSet &ps, const Lexer &lx) {
        rp_material m{};
        m.reflectance = constant(0.5f);
        m.eta = constant(1.5f);
        m.k = constant(0.0f);
        m.remap_roughness = ps.b("remaproughness", true) ? 1 : 0;
        float r = float(ps.f("roughness", 0));
        m.roughness_u = float(ps.f("uroughness", r));
        m.roughness_v = float(ps.f("vroughness", r));
        if (type == "diffuse") {
            m.kind = RP_MATERIAL_DIFFUSE;
            if (Param *p = ps.find("reflectance")) m.reflectance = albedo(p, lx);
        } else if (type == "conductor") {
            m.kind = RP_MATERIAL_CONDUCTOR;
            Param *eta = ps.find("eta");
            Param *k = ps.find("k");
            if (ps.find("reflectance")) lx.fail("conductor 'reflectance' parameterisation is not supported");
            if (!eta || !k) lx.fail("conductor needs explicit eta and k spectra");
            m.eta = unbounded(eta, lx);
            m.k = unbounded(k, lx);
        } else if (type == "dielectric") {
            m.kind = RP_MATERIAL_DIELECTRIC;
            if (Param *p = ps.find("eta")) m.eta = unbounded(p, lx);
        } else if (type == "interface") {
            m.kind = RP_MATERIAL_INTERFACE;
        } else {
            lx.fail("unsupported material '" + type + "'");
        }
        return m;
    }

    uint32_t current_material(const Lexer &lx) {
        if (g.material.empty()) return 0;
        auto it = named_materials.find(g.material);
        if (it == named_materials.end()) lx.fail("unknown material '" + g.material + "'");
        return it->second;
    }
    uint32_t medium_id(const std::string &name, const Lexer &lx) {
        if (name.empty()) return RP_NONE;
        auto it = named_media.find(name);
        if (it == named_media.end()) lx.fail("unknown medium '" + name + "'");
        return it->second;
    }

    void add_mesh(const std::vector<double> &P, const std::vector<uint32_t> &idx, const Lexer &lx) {
        if (P.size() % 3) lx.fail("P must hold xyz triples");
        uint32_t nv = uint32_t(P.size() / 3);
        if (idx.size() % 3) lx.fail("indices must hold triangles");
        rp_object obj{current_material(lx), RP_NONE, medium_id(g.inside, lx), medium_id(g.outside, lx)};
        if (g.area) {
            emissions.push_back(g.emission);
            obj.emission = uint32_t(emissions.size() - 1);
        }
        objects.push_back(obj);
        uint32_t oid = uint32_t(objects.size() - 1);
        uint32_t base = uint32_t(positions.size() / 3);
        for (uint32_t v = 0; v < nv; ++v) {
            auto w = xform_point(g.ctm, P[3 * v], P[3 * v + 1], P[3 * v + 2]);
            for (int a = 0; a < 3; ++a) positions.push_back(float(w[a]));
        }
        // A transform that flips handedness or ReverseOrientation flips the
        // geometric normal (PBRT): swap two vertices to keep emission sides.
        double det = g.ctm[0] * (g.ctm[5] * g.ctm[10] - g.ctm[6] * g.ctm[9]) -
                     g.ctm[1] * (g.ctm[4] * g.ctm[10] - g.ctm[6] * g.ctm[8]) +
                     g.ctm[2] * (g.ctm[4] * g.ctm[9] - g.ctm[5] * g.ctm[8]);
        bool flip = (det < 0) != g.reverse;
        for (size_t t = 0; t < idx.size(); t += 3) {
            for (int k = 0; k < 3; ++k)
                if (idx[t + k] >= nv) lx.fail("triangle index out of range");
            uint32_t a = base + idx[t], b = base + idx[t + 1], c = base + idx[t + 2];
            if (flip) std::swap(b, c);
            indices.insert(indices.end(), {a, b, c});
            tri_object.push_back(oid);
        }
    }

    void load_ply(const std::string &path, std::vector<double> &P, std::vector<uint32_t> &idx, const Lexer &lx) {
        std::ifstream in(path, std::ios::binary);
        if (!in) lx.fail("cannot open PLY '" + path + "'");
        std::string line, format;
        uint32_t nv = 0, nf = 0;
        std::vector<std::pair<std::string, std::string>> vprops; // (type, name)
        std::string face_count_type, face_index_type;
        std::string element;
        std::getline(in, line);
        if (line.rfind("ply", 0) != 0) lx.fail("not a PLY file: " + path);
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string kw;
            ls >> kw;
            if (kw == "format") ls >> format;
            else if (kw == "element") {
                uint32_t n;
                ls >> element >> n;
                if (element == "vertex") nv = n;
                else if (element == "face") nf = n;
                else if (n) lx.fail("PLY element '" + element + "' is not supported");
            } else if (kw == "property") {
                std::string t;
                ls >> t;
                if (t == "list") {
                    ls >> face_count_type >> face_index_type;
                } else {
                    std::string name;
                    ls >> name;
                    if (element == "vertex") vprops.push_back({t, name});
                }
            } else if (kw == "end_header") break;
        }
        auto size_of = [&](const std::string &t) -> int {
            if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
            if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
            if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
            if (t == "double" || t == "float64") return 8;
            lx.fail("PLY type '" + t + "' is not supported");
        };
        auto read_bin = [&](const std::string &t) -> double {
            unsigned char b[8] = {0};
            int n = size_of(t);
            in.read(reinterpret_cast<char *>(b), n);
            if (!in) lx.fail("truncated PLY " + path);
            if (t == "float" || t == "float32") {
                float f;
                std::memcpy(&f, b, 4);
                return f;
            }
            if (t == "double" || t == "float64") {
                double d;
                std::memcpy(&d, b, 8);
                return d;
            }
            uint64_t u = 0;
            for (int i = 0; i < n; ++i) u |= uint64_t(b[i]) << (8 * i);
            bool sign = t == "char" || t == "int8" || t == "short" || t == "int16" || t == "int" || t == "int32";
            if (sign && (u >> (8 * n - 1)) & 1) return double(int64_t(u) - (int64_t(1) << (8 * n)));
            return double(u);
        };
        bool ascii = format == "ascii";
        if (!ascii && format != "binary_little_endian") lx.fail("PLY format '" + format + "' is not supported");
        int ix = -1, iy = -1, iz = -1;
        for (size_t i = 0; i < vprops.size(); ++i) {
            if (vprops[i].second == "x") ix = int(i);
            if (vprops[i].second == "y") iy = int(i);
            if (vprops[i].second == "z") iz = int(i);
        }
        if (ix < 0 || iy < 0 || iz < 0) lx.fail("PLY without x/y/z");
        P.resize(3 * size_t(nv));
        for (uint32_t v = 0; v < nv; ++v) {
            std::vector<double> vals(vprops.size());
            for (size_t k = 0; k < vprops.size(); ++k) {
                if (ascii) in >> vals[k];
                else vals[k] = read_bin(vprops[k].first);
            }
            P[3 * v] = vals[ix], P[3 * v + 1] = vals[iy], P[3 * v + 2] = vals[iz];
        }
        for (uint32_t f = 0; f < nf; ++f) {
            uint32_t n;
            if (ascii) in >> n;
            else n = uint32_t(read_bin(face_count_type));
            std::vector<uint32_t> vi(n);
            for (uint32_t k = 0; k < n; ++k) {
                if (ascii) in >> vi[k];
                else vi[k] = uint32_t(read_bin(face_index_type));
            }
            for (uint32_t k = 1; k + 1 < n; ++k) idx.insert(idx.end(), {vi[0], vi[k], vi[k + 1]});
        }
        if (!in) lx.fail("truncated PLY " + path);
    }

    ParamSet params(Lexer &lx, Token &t, bool &have) {
        ParamSet ps;
        have = lx.next(t);
        while (have && t.kind == Token::string) {
            std::istringstream decl(t.text);
            Param p;
            decl >> p.type >> p.name;
            if (p.name.empty()) lx.fail("parameter declaration '" + t.text + "' needs a type and a name");
            if (p.type == "point" || p.type == "point3") p.type = "point3";
            if (p.type == "color") p.type = "rgb";
            if (!lx.next(t)) lx.fail("missing value for " + p.name);
            if (t.kind == Token::open) {
                while (lx.next(t) && t.kind != Token::close) {
                    if (t.kind == Token::number) p.nums.push_back(t.value);
                    else p.strs.push_back(t.text);
                }
            } else if (t.kind == Token::number) {
                p.nums.push_back(t.value);
            } else {
                p.strs.push_back(t.text);
            }
            ps.p.push_back(p);
            have = lx.next(t);
        }
        return ps;
    }

    static std::string read_file(const std::string &path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error("cannot open '" + path + "'");
        std::stringstream ss;
        ss << in.rdbuf();
        return ss.str();
    }

    void parse(const std::string &text, const std::string &file) {
        Lexer lx;
        lx.src = text;
        lx.file = file;
        Token t;
        bool have = lx.next(t);
        while (have) {
            if (t.kind != Token::word) lx.fail("expected a directive, found '" + t.text + "'");
            const std::string d = t.text;
            lx.directive_line = lx.token_line;
            auto nums = [&](int n) {
                std::vector<double> v;
                for (int i = 0; i < n; ++i) {
                    if (!lx.next(t) || t.kind != Token::number) lx.fail(d + " needs " + std::to_string(n) + " numbers");
                    v.push_back(t.value);
                }
                return v;
            };
            auto array16 = [&]() {
                std::vector<double> v;
                if (!lx.next(t) || t.kind != Token::open) lx.fail(d + " needs [ 16 numbers ]");
                while (lx.next(t) && t.kind != Token::close) v.push_back(t.value);
                if (v.size() != 16) lx.fail(d + " needs 16 numbers");
                M4 m;
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) m[i * 4 + j] = v[j * 4 + i]; // PBRT column-major
                return m;
            };
            auto word_arg = [&]() {
                if (!lx.next(t) || t.kind != Token::string) lx.fail(d + " needs a quoted name");
                return t.text;
            };
            if (d == "LookAt") {
                auto v = nums(9);
                double dx = v[3] - v[0], dy = v[4] - v[1], dz = v[5] - v[2];
                double dl = std::sqrt(dx * dx + dy * dy + dz * dz);
                dx /= dl, dy /= dl, dz /= dl;
                double ux = v[6], uy = v[7], uz = v[8], ul = std::sqrt(ux * ux + uy * uy + uz * uz);
                ux /= ul, uy /= ul, uz /= ul;
                // right = normalize(cross(up, dir)), newUp = cross(dir, right)
                double rx = uy * dz - uz * dy, ry = uz * dx - ux * dz, rz = ux * dy - uy * dx;
                double rl = std::sqrt(rx * rx + ry * ry + rz * rz);
                if (rl == 0) lx.fail("LookAt up vector parallel to the view direction");
                rx /= rl, ry /= rl, rz /= rl;
                double nx = dy * rz - dz * ry, ny = dz * rx - dx * rz, nz = dx * ry - dy * rx;
                M4 world_from_camera = {rx, nx, dx, v[0], ry, ny, dy, v[1], rz, nz, dz, v[2], 0, 0, 0, 1};
                M4 camera_from_world;
                if (!invert(world_from_camera, camera_from_world)) lx.fail("singular LookAt");
                g.ctm = mul(g.ctm, camera_from_world);
                have = lx.next(t);
            } else if (d == "Translate") {
                auto v = nums(3);
                g.ctm = mul(g.ctm, M4{1, 0, 0, v[0], 0, 1, 0, v[1], 0, 0, 1, v[2], 0, 0, 0, 1});
                have = lx.next(t);
            } else if (d == "Scale") {
                auto v = nums(3);
                g.ctm = mul(g.ctm, M4{v[0], 0, 0, 0, 0, v[1], 0, 0, 0, 0, v[2], 0, 0, 0, 0, 1});
                have = lx.next(t);
            } else if (d == "Rotate") {
                auto v = nums(4);
                double a = v[0] * 3.14159265358979323846 / 180, x = v[1], y = v[2], z = v[3];
                double l = std::sqrt(x * x + y * y + z * z);
                x /= l, y /= l, z /= l;
                double c = std::cos(a), s = std::sin(a);
                M4 r = {x * x + (1 - x * x) * c, x * y * (1 - c) - z * s, x * z * (1 - c) + y * s, 0,
                        x * y * (1 - c) + z * s, y * y + (1 - y * y) * c, y * z * (1 - c) - x * s, 0,
                        x * z * (1 - c) - y * s, y * z * (1 - c) + x * s, z * z + (1 - z * z) * c, 0,
                        0, 0, 0, 1};
                g.ctm = mul(g.ctm, r);
                have = lx.next(t);
            } else if (d == "Identity") {
                g.ctm = identity();
                have = lx.next(t);
            } else if (d == "Transform") {
                g.ctm = array16();
                have = lx.next(t);
            } else if (d == "ConcatTransform") {
                g.ctm = mul(g.ctm, array16());
                have = lx.next(t);
            } else if (d == "ReverseOrientation") {
                g.reverse = !g.reverse;
                have = lx.next(t);
            } else if (d == "Camera") {
                std::string type = word_arg();
                ParamSet ps = params(lx, t, have);
                if (type != "perspective") lx.fail("unsupported camera '" + type + "'");
                fov = ps.f("fov", 90);
                if (ps.f("lensradius", 0) != 0) lx.fail("thin-lens cameras are not supported");
                M4 world_from_camera;
                if (!invert(g.ctm, world_from_camera)) lx.fail("singular camera transform");
                {
                    // orthonormal axes (reflections allowed), as src/pbrt_render.cpp:543-557
                    const M4 &m = world_from_camera;
                    const double ax[3][3] = {{m[0], m[4], m[8]}, {m[1], m[5], m[9]}, {m[2], m[6], m[10]}};
                    auto dot3 = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
                    const double tol = 64 * 1.1920928955078125e-7;
                    for (int a = 0; a < 3; ++a)
                        for (int b2 = a; b2 < 3; ++b2)
                            if (std::fabs(dot3(ax[a], ax[b2]) - (a == b2 ? 1.0 : 0.0)) > tol)
                                lx.fail("Camera transform must be rigid");
                    if (m[12] != 0 || m[13] != 0 || m[14] != 0 || m[15] != 1) lx.fail("Camera transform must be rigid");
                }
                camera_world = world_from_camera;
                camera_medium_name = g.outside;
                have_camera = true;
            } else if (d == "Film") {
                word_arg();
                ParamSet ps = params(lx, t, have);
                info.width = uint32_t(ps.f("xresolution", 1280));
                info.height = uint32_t(ps.f("yresolution", 720));
                std::string fn = ps.s("filename", "");
                std::snprintf(info.output, sizeof info.output, "%s", fn.c_str());
            } else if (d == "PixelFilter") {
                std::string type = word_arg();
                ParamSet ps = params(lx, t, have);
                if (type != "box" || ps.f("xradius", 0.5) != 0.5 || ps.f("yradius", 0.5) != 0.5)
                    lx.fail("only the 0.5-radius box filter is supported");
            } else if (d == "Sampler") {
                std::string type = word_arg();
                ParamSet ps = params(lx, t, have);
                if (type != "sobol") lx.fail("unsupported sampler '" + type + "' (sobol only)");
                info.spp = uint32_t(ps.f("pixelsamples", 16));
                info.seed = uint64_t(uint32_t(int32_t(ps.f("seed", 0))));
                std::string r = ps.s("randomization", "fastowen");
                info.randomization = r == "none" ? RP_RANDOMIZE_NONE
                                     : r == "permutedigits" ? RP_RANDOMIZE_PERMUTE
                                     : r == "owen" ? RP_RANDOMIZE_OWEN
                                     : r == "fastowen" ? RP_RANDOMIZE_FASTOWEN
                                                       : ~0u;
                if (info.randomization == ~0u) lx.fail("unknown randomization '" + r + "'");
            } else if (d == "Integrator") {
                std::string type = word_arg();
                ParamSet ps = params(lx, t, have);
                if (type != "path" && type != "volpath") lx.fail("unsupported integrator '" + type + "'");
                info.volumetric = type == "volpath";
                info.max_depth = uint32_t(ps.f("maxdepth", 5));
                // PBRT's light sampler choice (default "bvh")
                const std::string ls = ps.s("lightsampler", "bvh");
                light_sampler = ls == "bvh" ? RP_LIGHT_SAMPLER_BVH
                                : ls == "power" ? RP_LIGHT_SAMPLER_POWER
                                : ls == "uniform" ? RP_LIGHT_SAMPLER_UNIFORM
                                                  : ~0u;
                if (light_sampler == ~0u) lx.fail("unknown lightsampler '" + ls + "'");
            } else if (d == "WorldBegin") {
                in_world = true;
                g.ctm = identity();
                have = lx.next(t);
            } else if (d == "AttributeBegin") {
                stack.push_back(g);
                have = lx.next(t);
            } else if (d == "AttributeEnd") {
                if (stack.empty()) lx.fail("unmatched AttributeEnd");
                g = stack.back();
                stack.pop_back();
                have = lx.next(t);
            } else if (d == "Material" || d == "MakeNamedMaterial") {
                std::string name = word_arg();
                ParamSet ps = params(lx, t, have);
                std::string type = d == "Material" ? name : ps.s("type", "");
                rp_material m = make_material(type, ps, lx);
                materials.push_back(m);
                std::string key = d == "Material" ? "\x01anon" + std::to_string(materials.size()) : name;
                named_materials[key] = uint32_t(materials.size() - 1);
                if (d == "Material") g.material = key;
            } else if (d == "NamedMaterial") {
                g.material = word_arg();
                if (!named_materials.count(g.material)) lx.fail("unknown material '" + g.material + "'");
                have = lx.next(t);
            } else if (d == "MakeNamedMedium") {
                std::string name = word_arg();
                ParamSet ps = params(lx, t, have);
                if (ps.s("type", "") != "homogeneous") lx.fail("only homogeneous media are supported");
                rp_medium md{};
                md.sigma_a = constant(1.0f);
                md.sigma_s = constant(1.0f);
                if (ps.find("preset")) lx.fail("medium presets are not supported");
                if (ps.find("Le")) lx.fail("emissive media are not supported");
                if (Param *p = ps.find("sigma_a")) md.sigma_a = unbounded(p, lx);
                if (Param *p = ps.find("sigma_s")) md.sigma_s = unbounded(p, lx);
                float sc = float(ps.f("scale", 1));
                md.sigma_a.scale *= sc;
                md.sigma_s.scale *= sc;
                md.g = float(ps.f("g", 0));
                media.push_back(md);
                named_media[name] = uint32_t(media.size() - 1);
            } else if (d == "MediumInterface") {
                g.inside = word_arg();
                have = lx.next(t);
                if (have && t.kind == Token::string) {
                    g.outside = t.text;
                    have = lx.next(t);
                } else {
                    g.outside = g.inside;
                }
            } else if (d == "AreaLightSource") {
                std::string type = word_arg();
                ParamSet ps = params(lx, t, have);
                if (type != "diffuse") lx.fail("unsupported area light '" + type + "'");
                rp_emission em{};
                if (ps.find("filename")) lx.fail("image area lights are not supported");
                if (ps.find("power")) lx.fail("area light 'power' is not supported");
                em.radiance = emission(ps.find("L"), ps.f("scale", 1), lx);
                em.two_sided = ps.b("twosided", false) ? 1 : 0;
                g.emission = em;
                g.area = true;
            } else if (d == "LightSource") {
                std::string type = word_arg();
                ParamSet ps = params(lx, t, have);
                if (type == "infinite") {
                    if (ps.find("filename")) lx.fail("image infinite lights are not supported");
                    desc.has_environment = 1;
                    if (ps.find("illuminance")) lx.fail("infinite light 'illuminance' is not supported");
                    desc.environment = emission(ps.find("L"), ps.f("scale", 1), lx);
                } else if (type == "spot") {
                    rp_spot_light sp{};
                    Param *from = ps.find("from");
                    Param *to = ps.find("to");
                    std::array<double, 3> f = {0, 0, 0}, tt = {0, 0, 1};
                    if (from) f = {from->nums.at(0), from->nums.at(1), from->nums.at(2)};
                    if (to) tt = {to->nums.at(0), to->nums.at(1), to->nums.at(2)};
                    auto pw = xform_point(g.ctm, f[0], f[1], f[2]);
                    auto tw = xform_point(g.ctm, tt[0], tt[1], tt[2]);
                    double ax = tw[0] - pw[0], ay = tw[1] - pw[1], az = tw[2] - pw[2];
                    double al = std::sqrt(ax * ax + ay * ay + az * az);
                    for (int a = 0; a < 3; ++a) sp.position[a] = float(pw[a]);
                    sp.direction[0] = float(ax / al), sp.direction[1] = float(ay / al), sp.direction[2] = float(az / al);
                    double cone = ps.f("coneangle", 30), delta = ps.f("conedelta", 5);
                    sp.cos_falloff_end = float(std::cos(cone * 3.14159265358979323846 / 180));
                    sp.cos_falloff_start = float(std::cos((cone - delta) * 3.14159265358979323846 / 180));
                    if (ps.find("power")) lx.fail("spot light 'power' is not supported");
                    sp.intensity = emission(ps.find("I"), ps.f("scale", 1), lx);
                    spots.push_back(sp);
                } else {
                    lx.fail("unsupported light '" + type + "'");
                }
            } else if (d == "Shape") {
                std::string type = word_arg();
                ParamSet ps = params(lx, t, have);
                std::vector<double> P;
                std::vector<uint32_t> idx;
                if (type == "trianglemesh") {
                    Param *pp = ps.find("P");
                    if (!pp) lx.fail("trianglemesh without P");
                    P = pp->nums;
                    if (Param *ip = ps.find("indices"))
                        for (double v : ip->nums) idx.push_back(uint32_t(v));
                    else if (P.size() == 9) idx = {0, 1, 2};
                    else lx.fail("trianglemesh without indices");
                } else if (type == "plymesh") {
                    std::string fn = ps.s("filename", "");
                    if (fn.empty()) lx.fail("plymesh without filename");
                    load_ply(fn[0] == '/' ? fn : base_dir + "/" + fn, P, idx, lx);
                } else {
                    lx.fail("unsupported shape '" + type + "'");
                }
                add_mesh(P, idx, lx);
            } else if (d == "Include" || d == "Import") {
                std::string fn = word_arg();
                if (++include_depth > 16) lx.fail("includes nested too deeply");
                std::string path = fn[0] == '/' ? fn : base_dir + "/" + fn;
                parse(read_file(path), path);
                --include_depth;
                have = lx.next(t);
            } else if (d == "WorldEnd") {
                have = lx.next(t);
            } else {
                lx.fail("unsupported directive '" + d + "'");
            }
        }
    }
    std::string camera_medium_name;

    void finish() {
        if (!have_camera) throw std::runtime_error("scene has no Camera");
        desc.api_version = RP_API_VERSION;
        // raster coefficients: screen window [-a, a] x [-1, 1] (a = aspect >= 1),
        // fov spans the shorter axis; d_cam = (sx t, sy t, 1)
        double w = info.width, h = info.height;
        double t = std::tan(fov * 0.5 * 3.14159265358979323846 / 180);
        double sx = w >= h ? t * w / h : t, sy = w >= h ? t : t * h / w;
        auto c = xform_vector(camera_world, -sx, sy, 1);
        auto dx = xform_vector(camera_world, 2 * sx, 0, 0);
        auto dy = xform_vector(camera_world, 0, -2 * sy, 0);
        auto o = xform_point(camera_world, 0, 0, 0);
        for (int a = 0; a < 3; ++a) {
            desc.camera.origin[a] = float(o[a]);
            desc.camera.corner[a] = float(c[a]);
            desc.camera.dx[a] = float(dx[a]);
            desc.camera.dy[a] = float(dy[a]);
        }
        desc.camera.width = info.width;
        desc.camera.height = info.height;
        desc.camera.medium = RP_NONE;
        if (!camera_medium_name.empty()) {
            auto it = named_media.find(camera_medium_name);
            if (it == named_media.end()) throw std::runtime_error("unknown camera medium '" + camera_medium_name + "'");
            desc.camera.medium = it->second;
        }
        desc.max_depth = info.max_depth;
        desc.hero_wavelengths = 0;
        desc.light_sampler = light_sampler;
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
        std::string err = validate_desc(desc);
        if (!err.empty()) throw std::runtime_error("translated scene is invalid: " + err);
    }
};

} // namespace
} // namespace rp

struct rp_pbrt {
    rp::Translator tr;
};

extern "C" {

int rp_pbrt_load_string(const char *text, const char *base_dir, rp_pbrt **out) {
    if (!text || !out) {
        rp::set_error("rp_pbrt_load: null argument");
        return RP_INVALID;
    }
    *out = nullptr;
    try {
        auto p = std::make_unique<rp_pbrt>();
        p->tr.base_dir = base_dir ? base_dir : ".";
        p->tr.parse(text, "<scene>");
        p->tr.finish();
        *out = p.release();
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        rp::set_error(std::string("rp_pbrt_load: ") + e.what());
        return RP_FAILURE;
    }
}

int rp_pbrt_load(const char *path, rp_pbrt **out) {
    if (!path || !out) {
        rp::set_error("rp_pbrt_load: null argument");
        return RP_INVALID;
    }
    *out = nullptr;
    try {
        std::string p(path);
        auto slash = p.find_last_of('/');
        std::string dir = slash == std::string::npos ? "." : p.substr(0, slash);
        auto scene = std::make_unique<rp_pbrt>();
        scene->tr.base_dir = dir;
        scene->tr.parse(rp::Translator::read_file(p), p);
        scene->tr.finish();
        *out = scene.release();
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        rp::set_error(std::string("rp_pbrt_load: ") + e.what());
        return RP_FAILURE;
    }
}

const rp_scene_desc *rp_pbrt_desc(const rp_pbrt *s) { return s ? &s->tr.desc : nullptr; }

int rp_pbrt_get_info(const rp_pbrt *s, rp_pbrt_info *info) {
    if (!s || !info) {
        rp::set_error("rp_pbrt_get_info: null argument");
        return RP_INVALID;
    }
    *info = s->tr.info;
    return RP_SUCCESS;
}

void rp_pbrt_destroy(rp_pbrt *s) { delete s; }

} // extern "C"
