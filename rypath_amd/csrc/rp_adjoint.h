// Path-replay adjoint (Vicini, Speierer, Jakob 2021) for the hero-wavelength
// path tracer.  Each path is traced twice from its Sobol sequence: the primal
// pass (the frame wavefront, which leaves every path's radiance and status
// in its per-lane arrays) gives the radiance L_k of every wavelength lane,
// the replay walks the same events with an AdjointHook and accumulates, for
// every factor a_v(theta)
// applied to the throughput, dLoss/dL_k * (radiance carried after v) *
// d log a_v / d theta.  Sampling decisions and MIS weights are detached (their
// densities are constants of the estimator), so nothing is stored per path.
//
// Parameters: binned diffuse reflectance tables, GGX conductor roughness
// (isotropic, PBRT remapping alpha = sqrt(r)) and binned medium sigma_s.
#pragma once
#include "rp_path.h"

namespace rpd {

constexpr uint32_t kMaxParams = 256;

struct GradTables {
    const int32_t *refl_base;   // per material: first parameter index, -1 = none
    const uint32_t *refl_bins;  // per material
    const int32_t *rough;       // per material: parameter index, -1 = none
    const int32_t *sigma_base;  // per medium
    const uint32_t *sigma_bins; // per medium
    uint32_t nparams;
    uint32_t sigma_medium;      // the medium whose sigma_s is differentiated (RP_NONE = none)
};

// Bin of a 471-sample table entry (nearest nanometre) for `bins` equal bins.
RP_DEV uint32_t lambda_bin(float lambda, uint32_t bins) {
    int i = int(floorf(lambda + 0.5f)) - 360;
    i = i < 0 ? 0 : (i > 470 ? 470 : i);
    uint32_t b = uint32_t(i) * bins / 471u;
    return b < bins ? b : bins - 1;
}

// d log f / d r for an isotropic rough conductor (alpha = sqrt(r) when remapped).
RP_DEV float conductor_dlogf_dr(const Lane &l, const rp_material &m, F3 wo, F3 wi) {
    if (l.smooth || !(wo.z * wi.z > 0)) return 0.0f;
    float a = l.ax;
    float r = m.roughness_u;
    float a_raw = m.remap_roughness ? __builtin_sqrtf(r) : r;
    if (a_raw < 1e-4f) return 0.0f; // clamped: alpha does not move with r
    F3 h = wo + wi;
    F3 mh = normalize(h);
    float s = mh.x * mh.x + mh.y * mh.y;
    float t = s / (a * a) + mh.z * mh.z;
    float dlogD = -2.0f / a + 4.0f * s / (a * a * a * t);
    float to = (wo.x * wo.x + wo.y * wo.y) / (wo.z * wo.z);
    float ti = (wi.x * wi.x + wi.y * wi.y) / (wi.z * wi.z);
    float so = __builtin_sqrtf(1.0f + a * a * to), si = __builtin_sqrtf(1.0f + a * a * ti);
    float lo = 0.5f * (so - 1.0f), li = 0.5f * (si - 1.0f);
    float dlo = 0.5f * a * to / so, dli = 0.5f * a * ti / si;
    float dlogf_da = dlogD - (dlo + dli) / (1.0f + lo + li);
    float da_dr = m.remap_roughness ? 0.5f / a : 1.0f;
    return dlogf_da * da_dr;
}

struct AdjointHook {
    static constexpr bool active = true;
    const GradTables *G;
    // LDS accumulator: parameter p at grad[p * stride].  With stride = block
    // size and grad = this thread's column, every thread owns its sums (no
    // atomics, bank-conflict free, a fixed summation order); with stride 0
    // the block shares one row through LDS atomics (many parameters).
    float *grad;
    uint32_t stride;
    float a[4];    // dLoss/dL_k
    float L[4];    // primal radiance
    float S[4];    // radiance accumulated so far in the replay
    // pending continuation factor (applied after the vertex's NEE)
    int32_t cont_idx[4];
    float cont_dlog[4];
    // pending NEE factor at the light direction
    int32_t nee_idx[4];
    float nee_dlog[4];

    RP_DEV void reset_pending() {
        for (int k = 0; k < 4; ++k) {
            cont_idx[k] = nee_idx[k] = -1;
            cont_dlog[k] = nee_dlog[k] = 0.0f;
        }
    }
    RP_DEV void accumulate(int32_t idx, float v) {
        if (!(idx >= 0 && v != 0.0f && __builtin_isfinite(v))) return;
        if (stride) grad[uint32_t(idx) * stride] += v;
        else atomicAdd(&grad[idx], v);
    }
    RP_DEV void add(const float *v) {
        for (int k = 0; k < 4; ++k) S[k] = S[k] + v[k];
    }
    // Free-flight scattering factor sigma_s * exp(-sigma_t t): d log / d sigma_s = 1/sigma_s - t.
    RP_DEV void medium_factor(uint32_t medium, const Wavelengths &wl, float t, const MediumCoeff &mc) {
        if (G->sigma_base[medium] < 0) return;
        for (int k = 0; k < 4; ++k) {
            if (!(mc.ss[k] > 0)) continue;
            int32_t idx = G->sigma_base[medium] + int32_t(lambda_bin(wl.lambda[k], G->sigma_bins[medium]));
            accumulate(idx, a[k] * (L[k] - S[k]) * (1.0f / mc.ss[k] - t));
        }
    }
    // Transmission factor exp(-sigma_t d): d log / d sigma_s = -d.
    RP_DEV void medium_pass(uint32_t medium, const Wavelengths &wl, float d) {
        if (G->sigma_base[medium] < 0) return;
        for (int k = 0; k < 4; ++k) {
            int32_t idx = G->sigma_base[medium] + int32_t(lambda_bin(wl.lambda[k], G->sigma_bins[medium]));
            accumulate(idx, a[k] * (L[k] - S[k]) * (-d));
        }
    }
    RP_DEV void factor(const rp_material &mat, uint32_t mid, const Lane *lanes, F3 wo, F3 wi, const Wavelengths &wl,
                       int32_t idx[4], float dlog[4]) {
        for (int k = 0; k < 4; ++k) {
            idx[k] = -1;
            dlog[k] = 0.0f;
        }
        if (mat.kind == RP_MATERIAL_DIFFUSE && G->refl_base[mid] >= 0) {
            for (int k = 0; k < 4; ++k) {
                if (!(lanes[k].r > 0) || !(wo.z * wi.z > 0)) continue;
                idx[k] = G->refl_base[mid] + int32_t(lambda_bin(wl.lambda[k], G->refl_bins[mid]));
                dlog[k] = 1.0f / lanes[k].r;
            }
        } else if (mat.kind == RP_MATERIAL_CONDUCTOR && G->rough[mid] >= 0) {
            for (int k = 0; k < 4; ++k) {
                idx[k] = G->rough[mid];
                dlog[k] = conductor_dlogf_dr(lanes[k], mat, wo, wi);
            }
        }
    }
    RP_DEV void continuation(const rp_material &mat, uint32_t mid, const Lane *lanes, F3 wo, F3 wi,
                             const Wavelengths &wl) {
        factor(mat, mid, lanes, wo, wi, wl, cont_idx, cont_dlog);
    }
    RP_DEV void nee(const rp_material &mat, uint32_t mid, const Lane *lanes, F3 wo, F3 wi, const Wavelengths &wl) {
        factor(mat, mid, lanes, wo, wi, wl, nee_idx, nee_dlog);
    }
    RP_DEV void nee_medium() {
        for (int k = 0; k < 4; ++k) {
            nee_idx[k] = -1;
            nee_dlog[k] = 0.0f;
        }
    }
};

// Replay of one path whose primal radiance is L (status `status`); its
// gradient terms go to hk's accumulator.
RP_DEV void adjoint_path(const Scene &S, const PathParams &P, const GradTables &G, Stack &st, uint32_t pixel,
                         uint64_t sample, uint32_t spp, const float *dl_dxyz, float4 L, uint32_t status,
                         AdjointHook &hk) {
    if (status != RP_PATH_SUCCESS) return;
    PathState ps;
    uint32_t w0, w1;
    path_begin(S, P, pixel, sample, ps, w0, w1);
    const Wavelengths wl = path_wavelengths(S, ps.w2);
    const float dx = dl_dxyz[3 * size_t(pixel)], dy = dl_dxyz[3 * size_t(pixel) + 1],
                dz = dl_dxyz[3 * size_t(pixel) + 2];
    const float Lp[4] = {L.x, L.y, L.z, L.w};
    for (int k = 0; k < 4; ++k) {
        hk.a[k] = 0.0f;
        if (wl.pdf[k] > 0) {
            int i = int(floorf(wl.lambda[k] + 0.5f)) - 360;
            i = i < 0 ? 0 : (i > 470 ? 470 : i);
            float4 c = S.cie[i];
            hk.a[k] = ((dx * c.x + dy * c.y + dz * c.z) / wl.pdf[k]) / (4.0f * float(spp));
        }
        hk.L[k] = Lp[k];
        hk.S[k] = 0.0f;
    }
    while (path_alive(S, ps)) {
        Hit h;
        traverse<false>(S, st, ps.o, ps.d, ps.tmax, h);
        hk.reset_pending();
        ShadowReq sr;
        bool cont = path_event(S, P, ps, h, sr, hk);
        if (sr.active) {
            float tr[4], mlen = 0.0f;
            bool vis = transmittance(S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr, P.boundary, &mlen,
                                     G.sigma_medium);
            if (vis) {
                float c[4];
                for (int k = 0; k < 4; ++k) c[k] = sr.val[k] * tr[k];
                if (!sr.delta_or_nomis)
                    for (int k = 0; k < 4; ++k) c[k] = c[k] * sr.w;
                hk.add(c);
                for (int k = 0; k < 4; ++k) {
                    hk.accumulate(hk.nee_idx[k], hk.a[k] * c[k] * hk.nee_dlog[k]);
                    if (G.sigma_medium != RP_NONE && mlen > 0) {
                        int32_t idx = G.sigma_base[G.sigma_medium] +
                                      int32_t(lambda_bin(wl.lambda[k], G.sigma_bins[G.sigma_medium]));
                        hk.accumulate(idx, hk.a[k] * c[k] * (-mlen));
                    }
                }
            }
        }
        for (int k = 0; k < 4; ++k) hk.accumulate(hk.cont_idx[k], hk.a[k] * (hk.L[k] - hk.S[k]) * hk.cont_dlog[k]);
        if (!cont) break;
    }
}

// Parameters up to which every thread keeps its own LDS column of sums
// (kPerThread): nparams * block * 4 B of dynamic LDS.
constexpr uint32_t kPerThreadParams = 96;

// Replay for paths [0, n) (path i: pixel i % npx, sample sample_begin +
// i / npx; primal radiance L_end[i], status end[i].x), grid-stride over a
// fixed grid so that block b always takes the same paths in the same order.
// Block b adds its sums to row b of `partial` ([gridDim.x][nparams]);
// k_grad_reduce folds the rows.  No global atomics, and with kPerThread no
// atomics at all: the gradient is the same bit for bit on every run.
template <int BS, bool kPerThread>
__global__ void __launch_bounds__(BS) k_adjoint(Scene S, PathParams P, GradTables G, uint32_t npx, uint64_t sample_begin,
                                                 uint32_t spp, uint64_t n, const float *dl_dxyz, const float4 *L_end,
                                                 const uint2 *end, float *partial) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    extern __shared__ float sgrad[]; // kPerThread: [nparams][BS]; else [nparams]
    const uint32_t np = G.nparams;
    if (kPerThread) {
        for (uint32_t p = 0; p < np; ++p) sgrad[p * BS + threadIdx.x] = 0.0f;
    } else {
        for (uint32_t p = threadIdx.x; p < np; p += BS) sgrad[p] = 0.0f;
        __syncthreads();
    }
    Stack st;
    st.init(s_node, s_t, BS);
    AdjointHook hk;
    hk.G = &G;
    hk.grad = kPerThread ? sgrad + threadIdx.x : sgrad;
    hk.stride = kPerThread ? BS : 0u;
    for (uint64_t lane = uint64_t(blockIdx.x) * BS + threadIdx.x; lane < n; lane += uint64_t(gridDim.x) * BS)
        adjoint_path(S, P, G, st, uint32_t(lane % npx), sample_begin + lane / npx, spp, dl_dxyz, L_end[lane],
                     end[lane].x, hk);
    __syncthreads();
    if (kPerThread) {
        // fold the columns: a fixed tree over the block's threads
        for (uint32_t h = BS / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h)
                for (uint32_t p = 0; p < np; ++p) sgrad[p * BS + threadIdx.x] += sgrad[p * BS + threadIdx.x + h];
            __syncthreads();
        }
        for (uint32_t p = threadIdx.x; p < np; p += BS) partial[size_t(blockIdx.x) * np + p] += sgrad[p * BS];
    } else {
        for (uint32_t p = threadIdx.x; p < np; p += BS) partial[size_t(blockIdx.x) * np + p] += sgrad[p];
    }
}

// grad[p] = sum over the rows of partial[rows][np], one block per parameter,
// a fixed summation order.
__global__ void __launch_bounds__(256) k_grad_reduce(const float *partial, uint32_t rows, uint32_t np, float *grad) {
    __shared__ float s[256];
    const uint32_t p = blockIdx.x;
    float acc = 0.0f;
    for (uint32_t r = threadIdx.x; r < rows; r += 256) acc += partial[size_t(r) * np + p];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[p] = s[0];
}

} // namespace rpd
