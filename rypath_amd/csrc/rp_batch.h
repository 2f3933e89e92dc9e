// Batched rendering of many scenes/views of uneven sizes in one
// device-resident work pool.  Two schedulers share the per-path arithmetic of
// rp_path.h, so every path (and every scene's film) is bit-identical to
// rendering that scene alone:
//  * persistent: one launch of a persistent megakernel; each wave takes 64
//    consecutive global path ids from a device counter until the pool is empty;
//  * wavefront: a fixed-capacity pool of compacted path state (two state sets,
//    as the frame wavefront): refill -> extend -> (sort) -> shade -> shadow ->
//    retire with device-side counts, grid-stride kernels and no host
//    synchronisation inside the bounce loop (replayable as one HIP graph).
//    New paths from any scene are appended behind the survivors each bounce.
#pragma once
#include "rp_wavefront.h"

namespace rpd {

struct BatchScene {
    Scene S;
    PathParams P;
    uint64_t path_begin; // first global path id of this scene
    uint64_t paths;      // npx * spp
    uint64_t sample_begin;
    float *xyz;          // planar per-path XYZ (3 * paths)
    uint32_t npx;
    uint32_t pad;
};

struct PoolCounters {
    uint32_t qn[2];     // live paths in state set 0 / 1
    uint32_t sn, en;    // shadow rays / ended paths of the current bounce
    uint64_t next;      // next global path id to issue
    uint64_t total;
    uint64_t done;      // retired paths
    unsigned long long rays[3]; // extension, shadow, failed
};

RP_DEV uint32_t find_scene(const BatchScene *scenes, uint32_t n, uint64_t g) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (scenes[mid].path_begin <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------- persistent megakernel
template <int BS>
__global__ void __launch_bounds__(BS) k_batch_persistent(const BatchScene *scenes, uint32_t nscenes,
                                                          unsigned long long *work, uint64_t total) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    Stack st;
    st.lds_node = s_node + threadIdx.x;
    st.lds_t = s_t + threadIdx.x;
    st.stride = BS;
    st.sp = 0;
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long ext = 0, sh = 0, failed = 0;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&work[0], 64ull);
        base = __shfl(base, 0);
        if (base >= total) break;
        uint64_t g = base + lane;
        if (g < total) {
            uint32_t s = find_scene(scenes, nscenes, g);
            const BatchScene &B = scenes[s];
            uint64_t local = g - B.path_begin;
            uint32_t pixel = uint32_t(local % B.npx);
            uint64_t sample = B.sample_begin + local / B.npx;
            PathResult r;
            trace_path(B.S, B.P, st, pixel, sample, r);
            float v[3] = {0, 0, 0};
            if (r.status == RP_PATH_SUCCESS) sensor_xyz(B.S, r.wl, r.L, v);
            else failed += 1;
            B.xyz[local] = v[0];
            B.xyz[B.paths + local] = v[1];
            B.xyz[2 * B.paths + local] = v[2];
            ext += r.ext_rays;
            sh += r.shadow_rays;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        ext += __shfl_down(ext, off);
        sh += __shfl_down(sh, off);
        failed += __shfl_down(failed, off);
    }
    if (lane == 0) {
        atomicAdd(&work[1], ext);
        atomicAdd(&work[2], sh);
        atomicAdd(&work[3], failed);
    }
}


// ------------------------------------------------------ wavefront pool
struct PoolView {
    WfState S[2];        // two compacted state sets (path-state fields only)
    uint32_t *sid[2];    // scene of each position, per state set
    FrameWf F;           // hits, shadow stream, sort order, histogram
    uint32_t *sh_sid;    // scene of each shadow ray
    uint4 *e_id;         // ended paths: (pixel, sample lo, sample hi, wavelength word)
    const BatchScene *scenes;
    uint32_t nscenes;
    uint32_t capacity;
    PoolCounters *c;
};
// Ended paths reuse the frame buffers' per-lane arrays: F.L_end (radiance)
// and F.end (status, scene).

#define RP_GRID_STRIDE(n)                                                                                     \
    for (uint32_t base_ = blockIdx.x * blockDim.x; base_ < (n); base_ += gridDim.x * blockDim.x)

__global__ void k_pool_init(PoolView V, uint64_t total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        V.c->qn[0] = V.c->qn[1] = 0;
        V.c->sn = V.c->en = 0;
        V.c->next = 0;
        V.c->total = total;
        V.c->done = 0;
        V.c->rays[0] = V.c->rays[1] = V.c->rays[2] = 0;
    }
}

RP_DEV uint32_t pool_new_paths(const PoolView &V, uint32_t p) {
    const uint32_t room = V.capacity - V.c->qn[p];
    const uint64_t remaining = V.c->total - V.c->next;
    return uint32_t(remaining < room ? remaining : room);
}

// New paths (in global id order, any scene) appended behind the survivors.
__global__ void k_pool_refill(PoolView V, uint32_t p) {
    const uint32_t n = pool_new_paths(V, p);
    const uint32_t live = V.c->qn[p];
    const uint64_t next = V.c->next;
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) {
            const uint64_t g = next + i;
            const uint32_t s = find_scene(V.scenes, V.nscenes, g);
            const BatchScene &BS = V.scenes[s];
            const uint64_t local = g - BS.path_begin;
            PathState ps;
            uint32_t w0, w1;
            path_begin(BS.S, BS.P, uint32_t(local % BS.npx), BS.sample_begin + local / BS.npx, ps, w0, w1);
            store_state(V.S[p], live + i, ps);
            V.sid[p][live + i] = s;
        }
    }
}
// Single block: publishes what k_pool_refill issued and clears the
// output-side counters and the sort histogram.
__global__ void k_pool_commit(PoolView V, uint32_t p, uint32_t bits) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t n = pool_new_paths(V, p);
        V.c->next += n;
        V.c->qn[p] += n;
        V.c->qn[1 - p] = 0;
        V.c->sn = 0;
        V.c->en = 0;
    }
    if (bits) {
        uint32_t K = 1u << bits;
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) {
            V.F.hist[k] = 0;
            V.F.hist[256 + k] = 0;
        }
    }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_pool_extend(PoolView V, uint32_t p) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    Stack st;
    st.lds_node = s_node + threadIdx.x;
    st.lds_t = s_t + threadIdx.x;
    st.stride = BS;
    st.sp = 0;
    const uint32_t n = V.c->qn[p];
    uint32_t rays = 0;
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) {
            const Scene &S = V.scenes[V.sid[p][i]].S;
            float4 a = V.S[p].o_t[i], b = V.S[p].d[i];
            Hit h;
            traverse<false>(S, st, F3{a.x, a.y, a.z}, F3{b.x, b.y, b.z}, a.w, h);
            V.F.hit_t[i] = make_float4(h.t, h.b0, h.b1, h.b2);
            V.F.hit_id[i] = make_uint2(h.prim, h.id);
            ++rays;
        }
    }
    wave_count(&V.c->rays[0], rays);
}

RP_DEV uint32_t pool_key(const PoolView &V, uint32_t p, uint32_t i, uint32_t bits) {
    return material_key(V.scenes[V.sid[p][i]].S, V.F.hit_id, i, bits);
}
__global__ void k_pool_sort_hist(PoolView V, uint32_t p, uint32_t bits) {
    __shared__ uint32_t h[256];
    const uint32_t K = 1u << bits, n = V.c->qn[p];
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
    __syncthreads();
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) atomicAdd(&h[pool_key(V, p, i, bits)], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
        if (h[k]) atomicAdd(&V.F.hist[k], h[k]);
}
// Each tile of blockDim entries is ranked in LDS.
__global__ void k_pool_sort_scatter(PoolView V, uint32_t p, uint32_t bits) {
    __shared__ uint32_t h[256], base[256];
    const uint32_t K = 1u << bits, n = V.c->qn[p];
    RP_GRID_STRIDE(n) {
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
        __syncthreads();
        uint32_t i = base_ + threadIdx.x, key = 0, rank = 0;
        if (i < n) {
            key = pool_key(V, p, i, bits);
            rank = atomicAdd(&h[key], 1u);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) base[k] = h[k] ? atomicAdd(&V.F.hist[256 + k], h[k]) : 0;
        __syncthreads();
        if (i < n) V.F.perm[base[key] + rank] = i;
        __syncthreads();
    }
}

// One path event per live path of set p: survivors are appended to set
// 1 - p, ended paths to the ended list, shadow rays to the shadow stream with
// their destination (survivor position, or kEnded | ended index).
template <int BS>
__global__ void __launch_bounds__(BS) k_pool_shade(PoolView V, uint32_t p, uint32_t sorted) {
    const uint32_t n = V.c->qn[p];
    const WfState &C = V.S[p], &N = V.S[1 - p];
    RP_GRID_STRIDE(n) {
        const uint32_t i = base_ + threadIdx.x;
        const bool valid = i < n;
        bool cont = false, shadow = false;
        uint32_t s = 0;
        PathState ps;
        ShadowReq sr;
        sr.active = false;
        if (valid) {
            const uint32_t j = sorted ? V.F.perm[i] : i;
            s = V.sid[p][j];
            const BatchScene &BSc = V.scenes[s];
            load_state(C, j, ps);
            float4 ht = V.F.hit_t[j];
            uint2 hi = V.F.hit_id[j];
            Hit h;
            h.prim = hi.x;
            h.id = hi.y;
            h.t = ht.x;
            h.b0 = ht.y;
            h.b1 = ht.z;
            h.b2 = ht.w;
            NoHook nh;
            cont = path_event(BSc.S, BSc.P, ps, h, sr, nh);
            shadow = sr.active;
            cont = cont && path_alive(BSc.S, ps);
        }
        const bool ended = valid && !cont;
        const uint32_t pos = wave_append(cont, &V.c->qn[1 - p]);
        if (cont) {
            store_state(N, pos, ps);
            V.sid[1 - p][pos] = s;
        }
        const uint32_t e = wave_append(ended, &V.c->en);
        if (ended) {
            V.F.L_end[e] = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
            V.F.end[e] = make_uint2(ps.status, s);
            V.e_id[e] = make_uint4(ps.pixel, uint32_t(ps.sample), uint32_t(ps.sample >> 32), ps.w2);
        }
        const uint32_t spos = wave_append(shadow, &V.c->sn);
        if (shadow) {
            V.F.sh_o[spos] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmax);
            V.F.sh_d[spos] = make_float4(sr.d.x, sr.d.y, sr.d.z, sr.w);
            V.F.sh_val[spos] = make_float4(sr.val[0], sr.val[1], sr.val[2], sr.val[3]);
            V.F.sh_meta[spos] = make_uint4(sr.medium, sr.delta_or_nomis ? 1u : 0u, ps.w2, cont ? pos : (kEnded | e));
            V.sh_sid[spos] = s;
        }
    }
}

// kInterfaces = false when no scene of the batch has null interfaces (the
// visibility test then needs no interface walk, and far fewer registers).
template <int BS, bool kInterfaces>
__global__ void __launch_bounds__(BS) k_pool_shadow(PoolView V, uint32_t p) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    Stack st;
    st.lds_node = s_node + threadIdx.x;
    st.lds_t = s_t + threadIdx.x;
    st.stride = BS;
    st.sp = 0;
    const uint32_t n = V.c->sn;
    const WfState &N = V.S[1 - p];
    uint32_t rays = 0;
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) {
            const BatchScene &BSc = V.scenes[V.sh_sid[i]];
            float4 o = V.F.sh_o[i], d = V.F.sh_d[i], v = V.F.sh_val[i];
            uint4 meta = V.F.sh_meta[i];
            ShadowReq sr;
            sr.o = {o.x, o.y, o.z};
            sr.tmax = o.w;
            sr.d = {d.x, d.y, d.z};
            sr.w = d.w;
            sr.val[0] = v.x, sr.val[1] = v.y, sr.val[2] = v.z, sr.val[3] = v.w;
            sr.medium = meta.x;
            sr.delta_or_nomis = meta.y != 0;
            Wavelengths wl = path_wavelengths(BSc.S, meta.z);
            float tr[4];
            bool vis = transmittance<kInterfaces>(BSc.S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr, BSc.P.boundary);
            if (vis) {
                float4 *dst = (meta.w & kEnded) ? &V.F.L_end[meta.w & ~kEnded] : &N.L[meta.w];
                float4 L = *dst;
                PathState ps;
                ps.L[0] = L.x, ps.L[1] = L.y, ps.L[2] = L.z, ps.L[3] = L.w;
                shadow_apply(ps, sr, vis, tr);
                *dst = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
            }
            ++rays;
        }
    }
    wave_count(&V.c->rays[1], rays);
}

// Ended paths: sensor XYZ into their scene's per-path planes.
__global__ void k_pool_retire(PoolView V) {
    const uint32_t n = V.c->en;
    uint32_t failed = 0, retired = 0;
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) {
            const uint2 em = V.F.end[i];
            const BatchScene &BSc = V.scenes[em.y];
            const float4 L = V.F.L_end[i];
            const uint4 id = V.e_id[i];
            Wavelengths wl = path_wavelengths(BSc.S, id.w);
            float Lv[4] = {L.x, L.y, L.z, L.w};
            float v[3] = {0, 0, 0};
            if (em.x == RP_PATH_SUCCESS) sensor_xyz(BSc.S, wl, Lv, v);
            else failed += 1;
            const uint64_t sample = uint64_t(id.y) | (uint64_t(id.z) << 32);
            const uint64_t local = (sample - BSc.sample_begin) * BSc.npx + id.x;
            BSc.xyz[local] = v[0];
            BSc.xyz[BSc.paths + local] = v[1];
            BSc.xyz[2 * BSc.paths + local] = v[2];
            retired += 1;
        }
    }
    unsigned long long r = retired;
    for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off);
    if ((threadIdx.x & 63) == 0 && r) atomicAdd(reinterpret_cast<unsigned long long *>(&V.c->done), r);
    wave_count(&V.c->rays[2], failed);
}

} // namespace rpd
