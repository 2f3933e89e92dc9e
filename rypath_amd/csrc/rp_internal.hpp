// Host-side internals of the rypath_amd library (not part of the C ABI).
#pragma once
#include "../../include/rp_api.h"
#include "../../include/rp_synthetic.h"

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace rp {

// ------------------------------------------------------------- errors
void set_error(const std::string &message);
const std::string &last_error();

// ------------------------------------------------------------- data files
// Directory holding sobol_tables.bin, cie_xyz_1931_2deg.csv, srgb_sigmoid.bin.
std::string data_dir();
struct SobolData {
    std::vector<uint32_t> directions; // [1024][52] (PBRT layout)
    std::vector<uint64_t> intervals;  // [25][52]
    std::vector<uint64_t> inverse;    // [26][52]
};
const SobolData &sobol_data();
const std::array<std::array<float, 3>, 471> &cie_xyz();
// PBRT stdillum-D65 at 360..830 nm (d65.bin), the sRGB colour space's illuminant.
const std::array<float, 471> &d65();
// sum_i s[i] * ybar[i] over the 471 one-nanometre samples (PBRT SpectrumToPhotometric).
double y_inner_product(const std::array<float, 471> &s);

// ------------------------------------------------------------- spectra
struct SigmoidCoefficients {
    bool constant = false;
    float value = 0;          // when constant
    std::array<float, 3> c{}; // descending powers of lambda (nm)
};
// PBRT v4 RGBToSpectrumTable lookup for linear sRGB in [0,1] (src/rgb_spectrum.cpp:62).
SigmoidCoefficients srgb_sigmoid(const std::array<float, 3> &rgb);
rp_spectrum spectrum_from_rgb(const std::array<float, 3> &rgb);
// Blackbody radiance normalised to its Wien peak, 471 samples (src/render.cpp:2185-2200).
std::array<float, 471> blackbody_table(double kelvin, double scale);
float spectrum_eval_host(const rp_spectrum &s, float lambda, const float *pool);

// ------------------------------------------------------------- BVH
// 64-byte quantised 4-wide node.  Child boxes are stored as 8-bit offsets from
// `origin` in units of 2^(e-127) per axis; the decoded box (origin + q*scale,
// evaluated in binary32 exactly as the traversal kernel does) always contains
// the true child box.
struct alignas(64) QNode {
    float origin[3];
    uint8_t exponent[3];
    uint8_t child_count;
    uint8_t qlo[3][4];
    uint8_t qhi[3][4];
    uint32_t child[4];  // internal: node index; leaf: kLeafBit | count << 27 | first (leaf order)
    uint8_t count[4];   // leaf triangle count (0 for internal children)
    uint32_t pad;
};
static_assert(sizeof(QNode) == 64, "QNode must be one 64-byte line");
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafCountShift = 27;
constexpr uint32_t kLeafFirstMask = (1u << 27) - 1;
constexpr uint32_t kEmptyChild = 0xffffffffu;

struct Bvh4 {
    std::vector<QNode> nodes;
    std::vector<uint32_t> tri_order; // leaf-order position -> scene triangle id
    uint32_t depth = 0;              // max node depth (root = 1)
    uint32_t max_stack = 0;          // worst-case traversal stack entries
    uint64_t leaf_count = 0;
    double seconds = 0;
};
Bvh4 build_bvh4(const float *positions, const uint32_t *indices, uint32_t triangle_count, uint32_t leaf_size,
                uint32_t bins, uint32_t threads);
inline float decode_lo(float origin, uint8_t e, uint8_t q) {
    float s;
    uint32_t bits = uint32_t(e) << 23;
    static_assert(sizeof(s) == 4);
    __builtin_memcpy(&s, &bits, 4);
    return origin + float(q) * s;
}

// ------------------------------------------------------------- lights
// One light: kind 0 triangle, 1 spot, 2 environment, 3 quad (p00 + u ex + v ey);
// index: scene triangle id (a quad's first triangle) or spot index; leaf: its
// light-BVH node; pmf / threshold / alias: the power sampler's alias table.
struct LightEntry {
    uint32_t kind, index, leaf, rect;
    float pmf, threshold;
    uint32_t alias;
    float inv_area;
    float p00[3], ex[3], ey[3];
    float ux[3], uy[3], uz[3]; // quad: unit edge directions and their cross product
    float lx, ly;              // quad: edge lengths
};
struct LightNodeHost {
    float low[3], high[3], dir[3];
    float cos_outer, cos_emission, power;
    uint32_t child, parent, light, flags;
};
struct LightTables {
    std::vector<LightEntry> lights;
    std::vector<LightNodeHost> nodes; // node i + 1 is the first child of interior node i
    uint32_t env_light = RP_NONE;
    float env_pmf = 0; // the environment's selection probability
};
// Light tables of a scene for its light_sampler; tri_light maps each
// triangle to its light (RP_NONE when not emissive).
LightTables build_light_tables(const rp_scene_desc &d, float scene_radius, std::vector<uint32_t> &tri_light);
float scene_bounding_radius(const rp_scene_desc &d);

// ------------------------------------------------------------- owned scene copy
struct HostScene {
    rp_scene_desc desc{};
    std::vector<float> positions, spectra;
    std::vector<uint32_t> indices, tri_object;
    std::vector<rp_object> objects;
    std::vector<rp_material> materials;
    std::vector<rp_emission> emissions;
    std::vector<rp_medium> media;
    std::vector<rp_spot_light> spots;
    void copy_from(const rp_scene_desc &d);
};
std::string validate_desc(const rp_scene_desc &d);

// SHA-256 over a canonical little-endian serialisation of every field of
// the description that changes rendered values.
std::array<uint8_t, 32> scene_identity(const rp_scene_desc &d);

} // namespace rp

struct rp_film;
namespace rp {
// Consistency of a running film (dimensions, provenance, finite sums).
std::string validate_film(const rp_film &f);
} // namespace rp

// Running raw-XYZ film (include/rp_io.h): the compensated per-pixel sums of
// the samples [0, samples), bound to one scene, seed, sampler and
// randomization, with the settings a checkpoint records beside it.
struct rp_film {
    uint32_t width = 0, height = 0;
    uint64_t samples = 0;
    uint64_t seed = 0;
    uint32_t randomization = 0;
    uint32_t sampler = RP_SAMPLER_SOBOL; // rp_sampler_kind of the samples held
    uint32_t model = 1;                  // film model: 0 demo (scalar demo transport), 1 PBRT
    bool packet = true;                  // unit-weight reconstruction sums (hero transport)
    uint64_t job_samples = 0, batch_size = 0; // the render job's settings (0 = unset)
    int32_t material_seed = 0;
    double exposure = 1.0, max_component = __builtin_inf();
    std::array<uint8_t, 32> identity{}; // scene identity (valid when samples > 0)
    std::vector<double> planes;         // 6 * width * height: sum X, corr X, sum Y, corr Y, sum Z, corr Z
    struct Segment {
        uint64_t begin, end;
        std::string build, device, architecture, os;
    };
    std::vector<Segment> segments; // provenance: sample ranges and what rendered them
    uint32_t format = 0;           // checkpoint format the film was read from (0: created here)
};
