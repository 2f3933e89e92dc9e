// Batched rendering of many scenes/views of uneven sizes in one
// device-resident work pool.  Two schedulers share the per-path arithmetic of
// rp_path.h, so every path (and every scene's film) is bit-identical to
// rendering that scene alone:
//  * persistent: one launch of a persistent megakernel; each wave takes 64
//    consecutive global path ids from a device counter until the pool is empty;
//  * wavefront: fixed-capacity slot pool; refill -> extend -> (sort) -> shade ->
//    shadow -> retire with device-side queue counts, grid-stride kernels and no
//    host synchronisation inside the bounce loop (replayable as one HIP graph).
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
    uint32_t qn[2];     // active queue sizes (ping-pong)
    uint32_t sn, rn;    // shadow / retire queue sizes
    uint32_t free_head, free_tail; // ring of free slots (monotonic counters)
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
    WfState B;
    const BatchScene *scenes;
    uint32_t nscenes;
    uint32_t capacity;
    uint32_t *sid;       // scene id per slot
    uint32_t *q[2];      // active queues
    uint32_t *qsorted;
    uint32_t *free_ring; // capacity entries
    uint32_t *retire;
    PoolCounters *c;
};

#define RP_GRID_STRIDE(n)                                                                                     \
    for (uint32_t base_ = blockIdx.x * blockDim.x; base_ < (n); base_ += gridDim.x * blockDim.x)

__global__ void k_pool_init(PoolView V, uint64_t total) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V.capacity) V.free_ring[i] = i;
    if (i == 0) {
        V.c->qn[0] = V.c->qn[1] = 0;
        V.c->sn = V.c->rn = 0;
        V.c->free_head = 0;
        V.c->free_tail = V.capacity;
        V.c->next = 0;
        V.c->total = total;
        V.c->done = 0;
        V.c->rays[0] = V.c->rays[1] = V.c->rays[2] = 0;
    }
}

// New paths into free slots, appended after the current input queue.
__global__ void k_pool_refill(PoolView V, uint32_t p) {
    const uint32_t avail = V.c->free_tail - V.c->free_head;
    const uint64_t remaining = V.c->total - V.c->next;
    const uint32_t n = uint32_t(remaining < avail ? remaining : avail);
    const uint32_t qbase = V.c->qn[p], head = V.c->free_head;
    const uint64_t next = V.c->next;
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) {
            uint32_t slot = V.free_ring[(head + i) % V.capacity];
            uint64_t g = next + i;
            uint32_t s = find_scene(V.scenes, V.nscenes, g);
            const BatchScene &BS = V.scenes[s];
            uint64_t local = g - BS.path_begin;
            PathState ps;
            uint32_t w0, w1;
            path_begin(BS.S, BS.P, uint32_t(local % BS.npx), BS.sample_begin + local / BS.npx, ps, w0, w1);
            store_state(V.B, slot, ps);
            V.sid[slot] = s;
            V.q[p][qbase + i] = slot;
        }
    }
}
// Single thread: publishes what k_pool_refill issued and clears the output-side counters.
__global__ void k_pool_commit(PoolView V, uint32_t p, uint32_t *hist, uint32_t bits) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t avail = V.c->free_tail - V.c->free_head;
        const uint64_t remaining = V.c->total - V.c->next;
        const uint32_t n = uint32_t(remaining < avail ? remaining : avail);
        V.c->free_head += n;
        V.c->next += n;
        V.c->qn[p] += n;
        V.c->qn[1 - p] = 0;
        V.c->sn = 0;
        V.c->rn = 0;
    }
    if (bits) {
        uint32_t K = 1u << bits;
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) {
            hist[k] = 0;
            hist[256 + k] = 0;
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
            uint32_t slot = V.q[p][i];
            const Scene &S = V.scenes[V.sid[slot]].S;
            float4 a = V.B.o_t[slot], b = V.B.d[slot];
            Hit h;
            traverse<false>(S, st, F3{a.x, a.y, a.z}, F3{b.x, b.y, b.z}, a.w, h);
            V.B.hit_t[slot] = make_float4(h.t, h.b0, h.b1, h.b2);
            V.B.hit_id[slot] = make_uint2(h.prim, h.id);
            ++rays;
        }
    }
    wave_count(&V.c->rays[0], rays);
}

RP_DEV uint32_t pool_key(const PoolView &V, uint32_t slot, uint32_t bits) {
    const Scene &S = V.scenes[V.sid[slot]].S;
    return material_key(S, V.B.hit_id, slot, bits);
}
__global__ void k_pool_sort_hist(PoolView V, uint32_t p, uint32_t bits, uint32_t *hist) {
    __shared__ uint32_t h[256];
    const uint32_t K = 1u << bits, n = V.c->qn[p];
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
    __syncthreads();
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) atomicAdd(&h[pool_key(V, V.q[p][i], bits)], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}
// One block per grid: each tile of blockDim entries is ranked in LDS.
__global__ void k_pool_sort_scatter(PoolView V, uint32_t p, uint32_t bits, uint32_t *hist) {
    __shared__ uint32_t h[256], base[256];
    const uint32_t K = 1u << bits, n = V.c->qn[p];
    RP_GRID_STRIDE(n) {
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
        __syncthreads();
        uint32_t i = base_ + threadIdx.x, key = 0, rank = 0, slot = 0;
        if (i < n) {
            slot = V.q[p][i];
            key = pool_key(V, slot, bits);
            rank = atomicAdd(&h[key], 1u);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) base[k] = h[k] ? atomicAdd(&hist[256 + k], h[k]) : 0;
        __syncthreads();
        if (i < n) V.qsorted[base[key] + rank] = slot;
        __syncthreads();
    }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_pool_shade(PoolView V, uint32_t p, uint32_t sorted) {
    const uint32_t n = V.c->qn[p];
    const uint32_t *q = sorted ? V.qsorted : V.q[p];
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        bool cont = false, shadow = false, retire = false;
        uint32_t slot = 0;
        if (i < n) {
            slot = q[i];
            const BatchScene &BSc = V.scenes[V.sid[slot]];
            PathState ps;
            load_state(V.B, slot, ps);
            float4 ht = V.B.hit_t[slot];
            uint2 hi = V.B.hit_id[slot];
            Hit h;
            h.prim = hi.x;
            h.id = hi.y;
            h.t = ht.x;
            h.b0 = ht.y;
            h.b1 = ht.z;
            h.b2 = ht.w;
            ShadowReq sr;
            NoHook nh;
            cont = path_event(BSc.S, BSc.P, ps, h, sr, nh);
            shadow = sr.active;
            cont = cont && path_alive(BSc.S, ps);
            retire = !cont;
            store_state<false>(V.B, slot, ps);
            if (shadow) {
                V.B.sh_o[slot] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmax);
                V.B.sh_d[slot] = make_float4(sr.d.x, sr.d.y, sr.d.z, sr.w);
                V.B.sh_val[slot] = make_float4(sr.val[0], sr.val[1], sr.val[2], sr.val[3]);
                V.B.sh_meta[slot] = make_uint2(sr.medium, sr.delta_or_nomis ? 1u : 0u);
            }
        }
        uint32_t pos = wave_append(cont, &V.c->qn[1 - p]);
        if (cont) V.q[1 - p][pos] = slot;
        uint32_t spos = wave_append(shadow, &V.c->sn);
        if (shadow) V.B.qs[spos] = slot;
        uint32_t rpos = wave_append(retire, &V.c->rn);
        if (retire) V.retire[rpos] = slot;
    }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_pool_shadow(PoolView V) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    Stack st;
    st.lds_node = s_node + threadIdx.x;
    st.lds_t = s_t + threadIdx.x;
    st.stride = BS;
    st.sp = 0;
    const uint32_t n = V.c->sn;
    uint32_t rays = 0;
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        if (i < n) {
            uint32_t slot = V.B.qs[i];
            const BatchScene &BSc = V.scenes[V.sid[slot]];
            float4 o = V.B.sh_o[slot], d = V.B.sh_d[slot], v = V.B.sh_val[slot];
            uint2 meta = V.B.sh_meta[slot];
            ShadowReq sr;
            sr.o = {o.x, o.y, o.z};
            sr.tmax = o.w;
            sr.d = {d.x, d.y, d.z};
            sr.w = d.w;
            sr.val[0] = v.x, sr.val[1] = v.y, sr.val[2] = v.z, sr.val[3] = v.w;
            sr.medium = meta.x;
            sr.delta_or_nomis = meta.y != 0;
            Wavelengths wl = path_wavelengths(BSc.S, V.B.id[slot].w);
            float tr[4];
            bool vis = transmittance(BSc.S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr, BSc.P.boundary);
            if (vis) {
                float4 L = V.B.L[slot];
                PathState ps;
                ps.L[0] = L.x, ps.L[1] = L.y, ps.L[2] = L.z, ps.L[3] = L.w;
                shadow_apply(ps, sr, vis, tr);
                V.B.L[slot] = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
            }
            ++rays;
        }
    }
    wave_count(&V.c->rays[1], rays);
}

// Finished paths: sensor XYZ into the scene's per-path planes, slot back to the free ring.
__global__ void k_pool_retire(PoolView V) {
    const uint32_t n = V.c->rn;
    uint32_t failed = 0, retired = 0;
    RP_GRID_STRIDE(n) {
        uint32_t i = base_ + threadIdx.x;
        bool ok = i < n;
        uint32_t slot = ok ? V.retire[i] : 0;
        if (ok) {
            const BatchScene &BSc = V.scenes[V.sid[slot]];
            float4 L = V.B.L[slot];
            uint4 id = V.B.id[slot];
            uint32_t status = V.B.meta2[slot].w;
            Wavelengths wl = path_wavelengths(BSc.S, id.w);
            float Lv[4] = {L.x, L.y, L.z, L.w};
            float v[3] = {0, 0, 0};
            if (status == RP_PATH_SUCCESS) sensor_xyz(BSc.S, wl, Lv, v);
            else failed += 1;
            uint64_t sample = uint64_t(id.y) | (uint64_t(id.z) << 32);
            uint64_t local = (sample - BSc.sample_begin) * BSc.npx + id.x;
            BSc.xyz[local] = v[0];
            BSc.xyz[BSc.paths + local] = v[1];
            BSc.xyz[2 * BSc.paths + local] = v[2];
            retired += 1;
        }
        uint32_t pos = wave_append(ok, &V.c->free_tail);
        if (ok) V.free_ring[pos % V.capacity] = slot;
    }
    unsigned long long r = retired;
    for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off);
    if ((threadIdx.x & 63) == 0 && r) atomicAdd(reinterpret_cast<unsigned long long *>(&V.c->done), r);
    wave_count(&V.c->rays[2], failed);
}

} // namespace rpd
