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
#include "rp_wavefront.h"

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
    // The block's LDS row of sums (np floats), added to with LDS atomics.
    // Measured against one LDS column per thread (np * 128 floats per block,
    // a fixed summation order): the columns limited the replay's shading
    // step to 2.5 waves/SIMD by LDS and took 8.47 ms per inverse step
    // against 8.17 ms with the shared row.
    float *grad;
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
        atomicAdd(&grad[idx], v);
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

// grad[p] += sum over the rows of partial[rows][np], one block per
// parameter, a fixed summation order.
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
    if (threadIdx.x == 0) grad[p] += s[0];
}

// ------------------------------------------------------ wavefront replay
// The replay runs as the frame wavefront does (generate -> extend -> shade
// -> shadow per bounce, compacted state), with the hook in the shading and
// shadow kernels:
//  * the hook's running radiance S is the path state's own L (path_event
//    adds every contribution to both, the shadow kernel adds NEE to L);
//  * a vertex's continuation term a*(L - S)*dlog needs S after that vertex's
//    NEE, so it rides with the survivor (cidx/cdlog, moved with compaction)
//    and is applied at the start of the next shading step;
//  * a path that ends has nothing after its last NEE: L - S = 0 there, so
//    its pending continuation contributes nothing;
//  * NEE and shadow-medium (sigma_s) terms are added in the shadow kernel
//    from the weights the shading step sent with the shadow ray.
// Paths whose primal status is not success contribute nothing, as in the
// primal film.
struct AdjWf {
    int4 *cidx[2];    // pending continuation parameter per lane k, per state set position
    float4 *cdlog[2]; // its d log / d theta
    float4 *sh_a;     // per shadow ray: dLoss/dL_k of its path
    int4 *sh_idx;     // NEE factor parameter per lane k
    float4 *sh_dlog;  // NEE factor d log / d theta
    const float4 *Lp; // primal radiance per lane
    const uint32_t *pstatus; // primal status per lane
    const float *dl;  // dLoss/dXYZ per pixel
    uint32_t spp;     // samples per pixel of the whole frame (a_k normalisation)
};

// a_k = dLoss/dL_k of a sample of `pixel` with wavelengths wl.
RP_DEV void adjoint_weights(const Scene &S, const AdjWf &A, uint32_t pixel, const Wavelengths &wl, float a[4]) {
    const float dx = A.dl[3 * size_t(pixel)], dy = A.dl[3 * size_t(pixel) + 1], dz = A.dl[3 * size_t(pixel) + 2];
    for (int k = 0; k < 4; ++k) {
        a[k] = 0.0f;
        if (wl.pdf[k] > 0) {
            int i = int(floorf(wl.lambda[k] + 0.5f)) - 360;
            i = i < 0 ? 0 : (i > 470 ? 470 : i);
            const float4 c = S.cie[i];
            a[k] = ((dx * c.x + dy * c.y + dz * c.z) / wl.pdf[k]) / (4.0f * float(A.spp));
        }
    }
}

// Zeroes the block's LDS row of np sums.
template <int BS> RP_DEV void clear_block_grad(float *sgrad, uint32_t np) {
    for (uint32_t q = threadIdx.x; q < np; q += BS) sgrad[q] = 0.0f;
    __syncthreads();
}
// Copies the block's LDS row into row blockIdx.x of `partial`.
template <int BS> RP_DEV void fold_block_grad(const float *sgrad, uint32_t np, float *partial) {
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < np; p += BS) partial[size_t(blockIdx.x) * np + p] = sgrad[p];
}

// One replay event for each path of set C[0, qn) (the frame wavefront's
// shading step plus the hook).  Dynamic LDS: np floats.
// Waves per SIMD requested from the compiler for k_adj_shade (register
// budget 512 / waves); 0 leaves the choice to the compiler.  Left alone it
// takes 169 VGPRs (2 waves/SIMD); 3 fits it in 168 without spills and
// measured 5.24 against 5.43 ms for the 256x256, 16 spp replay.
#ifndef RP_ADJ_SHADE_WAVES
#define RP_ADJ_SHADE_WAVES 3
#endif
#if RP_ADJ_SHADE_WAVES > 0
#define RP_ADJ_SHADE_ATTR __attribute__((amdgpu_waves_per_eu(RP_ADJ_SHADE_WAVES)))
#else
#define RP_ADJ_SHADE_ATTR
#endif

__global__ void __launch_bounds__(kWfBlock) RP_ADJ_SHADE_ATTR k_adj_shade(Scene S, PathParams P, GradTables G, AdjWf A, WfState C,
                                                         WfState N, FrameWf F, uint32_t p, uint32_t qn, uint32_t npx,
                                                         uint64_t sample_begin, float *partial) {
    __shared__ AppendLds<kWfBlock, 2> sh;
    extern __shared__ float sgrad[];
    const uint32_t np = G.nparams;
    clear_block_grad<kWfBlock>(sgrad, np);
    AdjointHook hk;
    hk.G = &G;
    hk.grad = sgrad;
    const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
    bool cont = false, shadow = false;
    PathState ps;
    ShadowReq sr;
    sr.active = false;
    if (i < qn) {
        load_state(C, i, ps);
        const uint32_t lane = uint32_t(ps.sample - sample_begin) * npx + ps.pixel;
        if (A.pstatus[lane] == RP_PATH_SUCCESS) {
            const float4 ht = F.hit_t[i];
            const uint2 hi = F.hit_id[i];
            Hit h;
            h.prim = hi.x;
            h.id = hi.y;
            h.t = ht.x;
            h.b0 = ht.y;
            h.b1 = ht.z;
            h.b2 = ht.w;
            const Wavelengths wl = path_wavelengths(S, ps.w2);
            adjoint_weights(S, A, ps.pixel, wl, hk.a);
            const float4 lp = A.Lp[lane];
            hk.L[0] = lp.x, hk.L[1] = lp.y, hk.L[2] = lp.z, hk.L[3] = lp.w;
            for (int k = 0; k < 4; ++k) hk.S[k] = ps.L[k];
            // the previous vertex's continuation, now that its NEE is in L
            const int4 ci = A.cidx[p][i];
            const float4 cd = A.cdlog[p][i];
            const int32_t cik[4] = {ci.x, ci.y, ci.z, ci.w};
            const float cdk[4] = {cd.x, cd.y, cd.z, cd.w};
            for (int k = 0; k < 4; ++k) hk.accumulate(cik[k], hk.a[k] * (hk.L[k] - hk.S[k]) * cdk[k]);
            hk.reset_pending();
            cont = path_event(S, P, ps, h, sr, hk);
            shadow = sr.active;
            cont = cont && path_alive(S, ps);
        }
    }
    const bool pred[2] = {cont, shadow};
    uint32_t *const queue[2] = {&F.qcount[0], &F.qcount[kQueueStride]};
    uint32_t at[2];
    block_append(sh, pred, queue, at);
    const uint32_t pos = at[0], spos = at[1];
    if (cont) {
        store_state(N, pos, ps); // with its identity: compaction moves the path
        A.cidx[1 - p][pos] = make_int4(hk.cont_idx[0], hk.cont_idx[1], hk.cont_idx[2], hk.cont_idx[3]);
        A.cdlog[1 - p][pos] = make_float4(hk.cont_dlog[0], hk.cont_dlog[1], hk.cont_dlog[2], hk.cont_dlog[3]);
    }
    if (shadow) {
        F.sh_o[spos] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmax);
        F.sh_d[spos] = make_float4(sr.d.x, sr.d.y, sr.d.z, sr.w);
        F.sh_val[spos] = make_float4(sr.val[0], sr.val[1], sr.val[2], sr.val[3]);
        F.sh_meta[spos] = make_uint4(sr.medium, sr.delta_or_nomis ? 1u : 0u, ps.w2, cont ? pos : kEnded);
        A.sh_a[spos] = make_float4(hk.a[0], hk.a[1], hk.a[2], hk.a[3]);
        A.sh_idx[spos] = make_int4(hk.nee_idx[0], hk.nee_idx[1], hk.nee_idx[2], hk.nee_idx[3]);
        A.sh_dlog[spos] = make_float4(hk.nee_dlog[0], hk.nee_dlog[1], hk.nee_dlog[2], hk.nee_dlog[3]);
    }
    fold_block_grad<kWfBlock>(sgrad, np, partial);
}

// Shadow rays of one replay bounce: the visible NEE contribution c goes to
// the survivor's L (as in the frame wavefront) and the NEE-factor and
// shadow-medium terms a_k * c_k * d log / d theta to the block's LDS row
// (shared, LDS atomics: at most eight terms per ray).
template <bool kInterfaces>
__global__ void __launch_bounds__(kWfBlock) k_adj_shadow(Scene S, PathParams P, GradTables G, AdjWf A, WfState N,
                                                          FrameWf F, uint32_t sn, float *partial) {
    __shared__ uint32_t s_node[kShortStack * kWfBlock];
    __shared__ float s_t[kShortStack * kWfBlock];
    __shared__ float srow[kMaxParams];
    const uint32_t np = G.nparams;
    for (uint32_t q = threadIdx.x; q < np; q += kWfBlock) srow[q] = 0.0f;
    __syncthreads();
    RP_STACK(st);
    st.init(s_node, s_t, kWfBlock);
    AdjointHook hk;
    hk.G = &G;
    hk.grad = srow;
    const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
    if (i < sn) {
        const float4 o = F.sh_o[i], d = F.sh_d[i], v = F.sh_val[i];
        const uint4 meta = F.sh_meta[i];
        ShadowReq sr;
        sr.o = {o.x, o.y, o.z};
        sr.tmax = o.w;
        sr.d = {d.x, d.y, d.z};
        sr.w = d.w;
        sr.val[0] = v.x, sr.val[1] = v.y, sr.val[2] = v.z, sr.val[3] = v.w;
        sr.medium = meta.x;
        sr.delta_or_nomis = meta.y != 0;
        const Wavelengths wl = path_wavelengths(S, meta.z);
        float tr[4], mlen = 0.0f;
        const bool vis = transmittance<kInterfaces>(S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr, P.boundary, &mlen,
                                                    G.sigma_medium);
        if (vis) {
            float c[4];
            for (int k = 0; k < 4; ++k) c[k] = sr.val[k] * tr[k];
            if (!sr.delta_or_nomis)
                for (int k = 0; k < 4; ++k) c[k] = c[k] * sr.w;
            if (!(meta.w & kEnded)) {
                const float4 L = N.L[meta.w];
                N.L[meta.w] = make_float4(L.x + c[0], L.y + c[1], L.z + c[2], L.w + c[3]);
            }
            const float4 a = A.sh_a[i], dg = A.sh_dlog[i];
            const int4 ix = A.sh_idx[i];
            const float ak[4] = {a.x, a.y, a.z, a.w}, dk[4] = {dg.x, dg.y, dg.z, dg.w};
            const int32_t ik[4] = {ix.x, ix.y, ix.z, ix.w};
            for (int k = 0; k < 4; ++k) {
                hk.accumulate(ik[k], ak[k] * c[k] * dk[k]);
                if (G.sigma_medium != RP_NONE && mlen > 0) {
                    const int32_t idx =
                        G.sigma_base[G.sigma_medium] + int32_t(lambda_bin(wl.lambda[k], G.sigma_bins[G.sigma_medium]));
                    hk.accumulate(idx, ak[k] * c[k] * (-mlen));
                }
            }
        }
    }
    fold_block_grad<kWfBlock>(srow, np, partial);
}

// Latency tail of the replay: one thread finishes each remaining path of
// C[0, qn), with the events of k_adj_shade and the shadow terms of
// k_adj_shadow in the same order per path (the pending continuation of a
// vertex is applied once its NEE is in L).  Gradient sums differ from the
// bounce kernels' only in float summation order.
template <bool kInterfaces>
__global__ void __launch_bounds__(kWfBlock) RP_PATH_ATTR k_adj_tail(Scene S, PathParams P, GradTables G, AdjWf A, WfState C,
                                                        uint32_t p, uint32_t qn, uint32_t npx, uint64_t sample_begin,
                                                        float *partial) {
    __shared__ uint32_t s_node[kShortStack * kWfBlock];
    __shared__ float s_t[kShortStack * kWfBlock];
    extern __shared__ float sgrad[];
    const uint32_t np = G.nparams;
    clear_block_grad<kWfBlock>(sgrad, np);
    RP_STACK(st);
    st.init(s_node, s_t, kWfBlock);
    AdjointHook hk;
    hk.G = &G;
    hk.grad = sgrad;
    const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
    if (i < qn) {
        PathState ps;
        load_state(C, i, ps);
        const uint32_t lane = uint32_t(ps.sample - sample_begin) * npx + ps.pixel;
        if (A.pstatus[lane] == RP_PATH_SUCCESS) {
            const Wavelengths wl = path_wavelengths(S, ps.w2);
            adjoint_weights(S, A, ps.pixel, wl, hk.a);
            const float4 lp = A.Lp[lane];
            hk.L[0] = lp.x, hk.L[1] = lp.y, hk.L[2] = lp.z, hk.L[3] = lp.w;
            const int4 ci = A.cidx[p][i];
            const float4 cd = A.cdlog[p][i];
            hk.cont_idx[0] = ci.x, hk.cont_idx[1] = ci.y, hk.cont_idx[2] = ci.z, hk.cont_idx[3] = ci.w;
            hk.cont_dlog[0] = cd.x, hk.cont_dlog[1] = cd.y, hk.cont_dlog[2] = cd.z, hk.cont_dlog[3] = cd.w;
            for (;;) {
                // the previous vertex's continuation, now that its NEE is in L
                for (int k = 0; k < 4; ++k) {
                    hk.S[k] = ps.L[k];
                    hk.accumulate(hk.cont_idx[k], hk.a[k] * (hk.L[k] - hk.S[k]) * hk.cont_dlog[k]);
                }
                Hit h;
                traverse<false>(S, st, ps.o, ps.d, ps.tmax, h);
                hk.reset_pending();
                ShadowReq sr;
                const bool cont = path_event(S, P, ps, h, sr, hk) && path_alive(S, ps);
                if (sr.active) {
                    float tr[4], mlen = 0.0f;
                    const bool vis = transmittance<kInterfaces>(S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr,
                                                                P.boundary, &mlen, G.sigma_medium);
                    if (vis) {
                        float c[4];
                        for (int k = 0; k < 4; ++k) c[k] = sr.val[k] * tr[k];
                        if (!sr.delta_or_nomis)
                            for (int k = 0; k < 4; ++k) c[k] = c[k] * sr.w;
                        for (int k = 0; k < 4; ++k) {
                            ps.L[k] = ps.L[k] + c[k];
                            hk.accumulate(hk.nee_idx[k], hk.a[k] * c[k] * hk.nee_dlog[k]);
                            if (G.sigma_medium != RP_NONE && mlen > 0) {
                                const int32_t idx = G.sigma_base[G.sigma_medium] +
                                                    int32_t(lambda_bin(wl.lambda[k], G.sigma_bins[G.sigma_medium]));
                                hk.accumulate(idx, hk.a[k] * c[k] * (-mlen));
                            }
                        }
                    }
                }
                if (!cont) break;
            }
        }
    }
    fold_block_grad<kWfBlock>(sgrad, np, partial);
}

// Scratch of the wavefront replay (one per device, beside the frame arena).
struct AdjBuffers {
    int4 *cidx[2] = {nullptr, nullptr};
    float4 *cdlog[2] = {nullptr, nullptr};
    float4 *sh_a = nullptr, *sh_dlog = nullptr, *Lp = nullptr;
    int4 *sh_idx = nullptr;
    uint32_t *pstatus = nullptr;
    float *partial = nullptr;
    uint64_t capacity = 0, partial_floats = 0;
    AdjBuffers() = default;
    AdjBuffers(const AdjBuffers &) = delete;
    AdjBuffers &operator=(const AdjBuffers &) = delete;
    ~AdjBuffers() { release(); }
    void release() {
        void *ptrs[] = {cidx[0], cidx[1], cdlog[0], cdlog[1], sh_a, sh_dlog, Lp, sh_idx, pstatus, partial};
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
        cidx[0] = cidx[1] = nullptr;
        cdlog[0] = cdlog[1] = nullptr;
        sh_a = sh_dlog = Lp = nullptr;
        sh_idx = nullptr;
        pstatus = nullptr;
        partial = nullptr;
        capacity = partial_floats = 0;
    }
    // partial: block rows of the largest launch plus room for the smaller
    // launches that follow it (up to 3 more of that size, 65536 rows at most)
    void ensure(uint64_t n, uint32_t np) {
        const uint64_t blocks = (n + kWfBlock - 1) / kWfBlock;
        const uint64_t pf = (blocks + std::min<uint64_t>(3 * blocks, 65536)) * std::max<uint32_t>(np, 1);
        if (n <= capacity && pf <= partial_floats) return;
        release();
        for (int s = 0; s < 2; ++s) wf_alloc(cidx[s], n), wf_alloc(cdlog[s], n);
        wf_alloc(sh_a, n), wf_alloc(sh_dlog, n), wf_alloc(sh_idx, n), wf_alloc(Lp, n), wf_alloc(pstatus, n);
        wf_alloc(partial, pf);
        capacity = n;
        partial_floats = pf;
    }
};

__global__ void __launch_bounds__(256) k_adj_prepare(const float4 *L_end, const uint2 *end, uint64_t n, float4 *Lp,
                                                     uint32_t *pstatus, int4 *cidx) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Lp[i] = L_end[i];
    pstatus[i] = end[i].x;
    cidx[i] = make_int4(-1, -1, -1, -1);
}

// Path replay of paths [0, n) (position i: pixel tile_pixel(i % npx), sample sample_begin +
// i / npx) whose primal radiance and status are in FB's per-lane arrays;
// adds dLoss/dtheta into grad (device, np floats).  Once at most `tail`
// paths are live, k_adj_tail finishes them.
inline void adjoint_replay(const Scene &S, const PathParams &P, const GradTables &G, FrameBuffers &FB,
                           AdjBuffers &AB, uint32_t npx, uint64_t sample_begin, uint64_t n, uint32_t spp,
                           const float *dl, float *grad, uint32_t tail, hipStream_t stream) {
    if (n > (uint64_t(1) << 31)) throw std::runtime_error("adjoint chunk too large");
    const uint32_t np = G.nparams;
    FB.ensure(n);
    AB.ensure(n, np);
    const FrameWf &F = FB.f;
    hipLaunchKernelGGL(k_adj_prepare, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream, F.L_end, F.end, n, AB.Lp,
                       AB.pstatus, AB.cidx[0]);
    int cur = 0;
    hipLaunchKernelGGL(k_wf_generate<false>, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream, S, P, npx, sample_begin, n,
                       FB.state[cur]);
    AdjWf A;
    A.cidx[0] = AB.cidx[0], A.cidx[1] = AB.cidx[1];
    A.cdlog[0] = AB.cdlog[0], A.cdlog[1] = AB.cdlog[1];
    A.sh_a = AB.sh_a, A.sh_idx = AB.sh_idx, A.sh_dlog = AB.sh_dlog;
    A.Lp = AB.Lp, A.pstatus = AB.pstatus, A.dl = dl, A.spp = spp;
    const size_t lds = size_t(np) * sizeof(float);
    // Block rows of successive launches go to consecutive rows of AB.partial;
    // one k_grad_reduce folds them into grad when the buffer is full and at
    // the end (instead of one reduction launch per kernel).
    const uint64_t row_cap = np ? AB.partial_floats / np : 0;
    uint64_t rows = 0;
    auto flush = [&] {
        if (np && rows) hipLaunchKernelGGL(k_grad_reduce, dim3(np), dim3(256), 0, stream, AB.partial, uint32_t(rows), np,
                                           grad);
        rows = 0;
    };
    auto rows_for = [&](uint32_t blocks) -> float * {
        if (rows + blocks > row_cap) flush();
        float *out = AB.partial + rows * np;
        rows += blocks;
        return out;
    };
    uint32_t qn = uint32_t(n);
    while (qn > 0) {
        const WfState &C = FB.state[cur], &N = FB.state[1 - cur];
        const dim3 grid((qn + kWfBlock - 1) / kWfBlock);
        if (qn <= tail) {
            float *part = rows_for(grid.x);
            if (S.has_interfaces)
                hipLaunchKernelGGL(k_adj_tail<true>, grid, dim3(kWfBlock), lds, stream, S, P, G, A, C, uint32_t(cur),
                                   qn, npx, sample_begin, part);
            else
                hipLaunchKernelGGL(k_adj_tail<false>, grid, dim3(kWfBlock), lds, stream, S, P, G, A, C, uint32_t(cur),
                                   qn, npx, sample_begin, part);
            break;
        }
        hipLaunchKernelGGL((k_wf_extend<false, false>), grid, dim3(kWfBlock), 0, stream, S, C, F, qn, nullptr);
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(256), 0, stream, F.qcount, F.hist, 0u);
        float *part = rows_for(grid.x);
        hipLaunchKernelGGL(k_adj_shade, grid, dim3(kWfBlock), lds, stream, S, P, G, A, C, N, F, uint32_t(cur), qn, npx,
                           sample_begin, part);
        WF_CHECK(hipMemcpyAsync(FB.host_count, F.qcount, (kQueueStride + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                stream));
        WF_CHECK(hipStreamSynchronize(stream));
        const uint32_t next = FB.host_count[0], sn = FB.host_count[kQueueStride];
        if (next > qn || sn > qn) throw std::runtime_error("adjoint queue counts out of range");
        if (sn) {
            const dim3 sg((sn + kWfBlock - 1) / kWfBlock);
            float *spart = rows_for(sg.x);
            if (S.has_interfaces)
                hipLaunchKernelGGL(k_adj_shadow<true>, sg, dim3(kWfBlock), 0, stream, S, P, G, A, N, F, sn, spart);
            else
                hipLaunchKernelGGL(k_adj_shadow<false>, sg, dim3(kWfBlock), 0, stream, S, P, G, A, N, F, sn, spart);
        }
        cur = 1 - cur;
        qn = next;
    }
    flush();
    WF_CHECK(hipGetLastError());
}

} // namespace rpd
