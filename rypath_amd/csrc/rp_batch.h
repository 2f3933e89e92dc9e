// Batched rendering of many scenes/views of uneven sizes in one
// device-resident work pool.  Two schedulers share the per-path arithmetic of
// rp_path.h, so every path (and every scene's film) is bit-identical to
// rendering that scene alone:
//  * persistent: one launch of a persistent megakernel; each wave takes 64
//    consecutive global path ids from a device counter until the pool is empty;
//  * wavefront: a fixed-capacity pool of compacted path state (two state sets,
//    as the frame wavefront): refill -> extend -> (sort) -> shade -> shadow ->
//    retire with device-side counts and no host synchronisation inside the
//    bounce loop (replayable as one HIP graph).  New paths from any scene are
//    appended behind the survivors each bounce; the heavy kernels run one
//    scene per wave (scene record in SGPRs), with per-lane kernels for the
//    rare waves that hold several scenes.
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

// The three queues the shading kernel appends to (survivors, ended paths,
// shadow rays) each have their own 256-B line: atomics on one line are
// serialised (kCounterStride).
struct PoolCounters {
    uint32_t qn[2];     // live paths in state set 0 / 1
    uint32_t ov, ovs;   // lanes of mixed-scene waves handed over: path events / shadow rays
    uint64_t next;      // next global path id to issue
    uint64_t total;
    uint64_t done;      // retired paths
    unsigned long long rays[3]; // extension, shadow, failed
    unsigned long long mixed;   // path events and shadow rays handled per lane
    alignas(256) uint32_t sn;   // shadow rays of the current bounce
    alignas(256) uint32_t en;   // ended paths of the current bounce
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

// Runs f(scene) once per distinct scene among the wave's active lanes
// (a waterfall over readfirstlane).  The scene index is wave-uniform inside
// f and the table is read through the constant address space, so the scene's
// fields load through the scalar unit into SGPRs exactly as a kernel
// argument's do; a per-lane scene would hold every node, triangle and
// material base address in VGPRs (190 instead of 160 in the shade kernel:
// 2 waves per SIMD instead of 3).  The loop is wave-uniform (a lane never
// leaves it early, so the compiler cannot move f past the exit where u is
// per-lane again), and lane membership is taken from a ballot rather than
// from `s == u`, which the compiler would use to substitute the per-lane s
// for u inside f.  Refill appends paths in scene order, so a wave nearly
// always holds a single scene and f runs once.
typedef const __attribute__((address_space(4))) BatchScene ConstBatchScene;
template <class Fn>
RP_DEV void for_scene(const BatchScene *scenes, uint32_t s, Fn &&f) {
    const uint32_t lane = __lane_id();
    uint64_t pending = __ballot(true);
    while (pending) {
        const uint32_t u = __builtin_amdgcn_readlane(s, __builtin_ctzll(pending));
        const uint64_t mine = __ballot(s == u);
        if ((mine >> lane) & 1) {
            ConstBatchScene *t = (ConstBatchScene *)scenes;
            f((const BatchScene &)t[u]);
        }
        pending &= ~mine;
    }
}

// ---------------------------------------------------- persistent megakernel
template <int BS>
__global__ void __launch_bounds__(BS) RP_PATH_ATTR k_batch_persistent(const BatchScene *scenes, uint32_t nscenes,
                                                          unsigned long long *work, uint64_t total) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    RP_STACK(st);
    st.init(s_node, s_t, BS);
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long ext = 0, sh = 0, failed = 0;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&work[0], 64ull);
        base = __shfl(base, 0);
        if (base >= total) break;
        uint64_t g = base + lane;
        if (g < total) {
            // a wave's 64 consecutive ids span one scene, or two at a boundary
            for_scene(scenes, find_scene(scenes, nscenes, g), [&](const BatchScene &B) {
                const uint64_t local = g - B.path_begin, s = local / B.npx;
                const uint32_t pixel = tile_pixel(B.S, uint32_t(local % B.npx));
                PathResult r;
                trace_path(B.S, B.P, st, pixel, B.sample_begin + s, r);
                float v[3] = {0, 0, 0};
                if (r.status == RP_PATH_SUCCESS) sensor_xyz(B.S, r.wl, r.L, v);
                else failed += 1;
                const uint64_t lane = s * B.npx + pixel;
                B.xyz[lane] = v[0];
                B.xyz[B.paths + lane] = v[1];
                B.xyz[2 * B.paths + lane] = v[2];
                ext += r.ext_rays;
                sh += r.shadow_rays;
            });
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        ext += __shfl_down(ext, off);
        sh += __shfl_down(sh, off);
        failed += __shfl_down(failed, off);
    }
    if (lane == 0) {
        atomicAdd(&work[kCounterStride], ext);
        atomicAdd(&work[2 * kCounterStride], sh);
        if (failed) atomicAdd(&work[3 * kCounterStride], failed);
    }
}


// ------------------------------------------------------ wavefront pool
struct PoolView {
    WfState S[2];        // two compacted state sets (path-state fields only)
    uint32_t *sid[2];    // scene of each position, per state set (| kSidFresh: new, compact)
    FrameWf F;           // hits, shadow stream, sort order, histogram
    uint32_t *sh_sid;    // scene of each shadow ray
    uint4 *e_id;         // ended paths: (pixel, sample lo, sample hi, wavelength word)
    uint32_t *ov;        // positions handed from scene-uniform to per-lane kernels
    const BatchScene *scenes;
    uint32_t nscenes;
    uint32_t capacity;
    PoolCounters *c;
};
// Ended paths reuse the frame buffers' per-lane arrays: F.L_end (radiance)
// and F.end (status, scene).
// A path refilled this iteration carries this bit in its sid entry: its
// state is compact (k_pool_refill) and the extend and shading kernels
// rebuild the rest.  The flag travels with the scene index the kernels read
// anyway, so no kernel waits on a separate load before its state loads.
constexpr uint32_t kSidFresh = 0x80000000u;

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
        V.c->mixed = 0;
    }
}

RP_DEV uint32_t pool_new_paths(const PoolView &V, uint32_t p) {
    const uint32_t room = V.capacity - V.c->qn[p];
    const uint64_t remaining = V.c->total - V.c->next;
    return uint32_t(remaining < room ? remaining : room);
}

// New paths (in global id order, any scene) appended behind the survivors,
// compactly (direction and wavelength word, Sobol index; as
// k_wf_generate<true>): the extend and shading kernels rebuild the rest.
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
            path_begin(BS.S, BS.P, tile_pixel(BS.S, uint32_t(local % BS.npx)), BS.sample_begin + local / BS.npx, ps,
                       w0, w1);
            store_compact(V.S[p], live + i, ps, BS.sample_begin);
            V.sid[p][live + i] = s | kSidFresh;
        }
    }
}
// Single block: publishes what k_pool_refill issued, counts this
// iteration's extension rays (one per live path) and clears the output-side
// counters and the sort histogram.
__global__ void k_pool_commit(PoolView V, uint32_t p, uint32_t bits) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t n = pool_new_paths(V, p);
        V.c->next += n;
        V.c->qn[p] += n;
        V.c->rays[0] += V.c->qn[p];
        V.c->qn[1 - p] = 0;
        V.c->sn = 0;
        V.c->en = 0;
        V.c->ov = V.c->ovs = 0;
    }
    if (bits) {
        uint32_t K = 1u << bits;
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) {
            V.F.hist[k * kQueueStride] = 0;
            V.F.hist[(256 + k) * kQueueStride] = 0;
        }
    }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_pool_extend(PoolView V, uint32_t p) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    RP_STACK(st);
    st.init(s_node, s_t, BS);
    const uint32_t n = V.c->qn[p], nb = (n + BS - 1) / BS;
    if (blockIdx.x >= nb) return;
    const uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i < n) {
        const uint32_t sw = V.sid[p][i];
        const bool fresh = (sw & kSidFresh) != 0; // a new path leaves its scene's camera
        const float4 b = V.S[p].d[i];
        float4 a = fresh ? make_float4(0.0f, 0.0f, 0.0f, __builtin_inff()) : V.S[p].o_t[i];
        Hit h;
        for_scene(V.scenes, sw & ~kSidFresh, [&](const BatchScene &B) {
            if (fresh) a = make_float4(B.S.camera.origin[0], B.S.camera.origin[1], B.S.camera.origin[2], a.w);
            traverse<false>(B.S, st, F3{a.x, a.y, a.z}, F3{b.x, b.y, b.z}, a.w, h);
        });
        V.F.hit_t[i] = make_float4(h.t, h.b0, h.b1, h.b2);
        V.F.hit_id[i] = make_uint2(h.prim, h.id);
    }
}

RP_DEV uint32_t pool_key(const PoolView &V, uint32_t p, uint32_t i, uint32_t bits) {
    return material_key(V.scenes[V.sid[p][i] & ~kSidFresh].S, V.F.hit_id, i, bits);
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
        if (h[k]) atomicAdd(&V.F.hist[k * kQueueStride], h[k]);
}
// Each tile of kSortBlockItems entries (16 per thread, coalesced) is ranked
// in LDS and reserves its key ranges with one global atomic per key: with
// 256-entry tiles those serialised atomics dominated the sort.
__global__ void k_pool_sort_scatter(PoolView V, uint32_t p, uint32_t bits) {
    __shared__ uint32_t h[256], base[256];
    const uint32_t K = 1u << bits, n = V.c->qn[p];
    for (uint32_t tile = blockIdx.x * kSortBlockItems; tile < n; tile += gridDim.x * kSortBlockItems) {
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
        __syncthreads();
        uint32_t packed[kSortItems]; // key << 12 | rank within the tile (< 4096)
#pragma unroll
        for (uint32_t j = 0; j < kSortItems; ++j) {
            const uint32_t i = tile + j * 256 + threadIdx.x;
            packed[j] = 0;
            if (i < n) {
                const uint32_t key = pool_key(V, p, i, bits);
                packed[j] = key << 12 | atomicAdd(&h[key], 1u);
            }
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
            base[k] = h[k] ? atomicAdd(&V.F.hist[(256 + k) * kQueueStride], h[k]) : 0;
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kSortItems; ++j) {
            const uint32_t i = tile + j * 256 + threadIdx.x;
            if (i < n) V.F.perm[base[packed[j] >> 12] + (packed[j] & 4095u)] = i;
        }
        __syncthreads();
    }
}

// Scene-uniform waves.  The shading and shadow kernels run one scene per
// wave with the scene's fields in SGPRs (as for_scene does, but without a
// loop: the path state held across a waterfall loop costs the shade kernel
// about 40 VGPRs).  A wave whose valid lanes hold more than one scene keeps
// the lanes of its first valid lane's scene and hands the others, through
// the `ov` list, to a per-lane kernel (k_pool_*_mixed) launched right after
// it.  Such waves are rare: refill issues paths in scene order, so a wave
// mixes scenes only where that order changes.
RP_DEV uint32_t wave_scene(bool valid, uint32_t s, bool &mine) {
    const uint64_t vmask = __ballot(valid);
    const uint32_t u = __builtin_amdgcn_readlane(s, vmask ? __builtin_ctzll(vmask) : 0);
    mine = (__ballot(valid && s == u) >> __lane_id()) & 1;
    return u;
}
RP_DEV const BatchScene &uniform_scene(const BatchScene *scenes, uint32_t u) {
    return (const BatchScene &)((ConstBatchScene *)scenes)[u];
}

// One path event for the path at S[p][j] of scene s (B = its record) when
// `valid`: survivors are appended to set 1 - p, ended paths to the ended
// list, shadow rays to the shadow stream with their destination (survivor
// position, or kEnded | ended index).  Every thread of the block calls this
// together (the appends are block-aggregated).
template <int BS>
RP_DEV void pool_shade_path(const PoolView &V, uint32_t p, bool valid, uint32_t j, uint32_t s, bool fresh,
                            const BatchScene &B, AppendLds<BS, 3> &sh) {
    const WfState &C = V.S[p], &N = V.S[1 - p];
    bool cont = false, shadow = false;
    PathState ps;
    ShadowReq sr;
    sr.active = false;
    if (valid) {
        if (fresh) load_compact(B.S, B.P, C, j, B.sample_begin, ps);
        else load_state(C, j, ps);
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
        cont = path_event(B.S, B.P, ps, h, sr, nh);
        shadow = sr.active;
        cont = cont && path_alive(B.S, ps);
    }
    const bool ended = valid && !cont;
    const bool pred[3] = {cont, ended, shadow};
    uint32_t *const queue[3] = {&V.c->qn[1 - p], &V.c->en, &V.c->sn};
    uint32_t at[3];
    block_append(sh, pred, queue, at);
    const uint32_t pos = at[0], e = at[1], spos = at[2];
    if (cont) {
        store_state(N, pos, ps);
        V.sid[1 - p][pos] = s;
    }
    if (ended) {
        V.F.L_end[e] = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
        V.F.end[e] = make_uint2(ps.status, s);
        V.e_id[e] = make_uint4(ps.pixel, uint32_t(ps.sample), uint32_t(ps.sample >> 32), ps.w2);
    }
    if (shadow) {
        V.F.sh_o[spos] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmax);
        V.F.sh_d[spos] = make_float4(sr.d.x, sr.d.y, sr.d.z, sr.w);
        V.F.sh_val[spos] = make_float4(sr.val[0], sr.val[1], sr.val[2], sr.val[3]);
        V.F.sh_meta[spos] = make_uint4(sr.medium, sr.delta_or_nomis ? 1u : 0u, ps.w2, cont ? pos : (kEnded | e));
        V.sh_sid[spos] = s;
    }
}

// One thread per live path of set p (the grid covers the capacity; blocks
// past the live count return at once).
template <int BS>
__global__ void __launch_bounds__(BS) k_pool_shade(PoolView V, uint32_t p, uint32_t sorted) {
    __shared__ AppendLds<BS, 3> sh;
    const uint32_t n = V.c->qn[p], nb = (n + BS - 1) / BS;
    if (blockIdx.x >= nb) return;
    const uint32_t i = blockIdx.x * BS + threadIdx.x;
    const bool valid = i < n;
    uint32_t j = 0, s = 0;
    bool fresh = false;
    if (valid) {
        j = sorted ? V.F.perm[i] : i;
        const uint32_t sw = V.sid[p][j];
        s = sw & ~kSidFresh;
        fresh = (sw & kSidFresh) != 0;
    }
    bool mine;
    const uint32_t u = wave_scene(valid, s, mine);
    const bool other = valid && !mine;
    const uint32_t o = wave_append(other, &V.c->ov);
    if (other) V.ov[o] = j;
    pool_shade_path<BS>(V, p, mine, j, u, fresh, uniform_scene(V.scenes, u), sh);
}

// The path events k_pool_shade handed over, one scene per lane.
__global__ void __launch_bounds__(256) k_pool_shade_mixed(PoolView V, uint32_t p) {
    __shared__ AppendLds<256, 3> sh;
    const uint32_t n = V.c->ov;
    uint32_t mixed = 0;
    RP_GRID_STRIDE(n) {
        const uint32_t i = base_ + threadIdx.x;
        const bool valid = i < n;
        const uint32_t j = valid ? V.ov[i] : 0u;
        const uint32_t sw = valid ? V.sid[p][j] : 0u, s = sw & ~kSidFresh;
        pool_shade_path<256>(V, p, valid, j, s, (sw & kSidFresh) != 0, V.scenes[s], sh);
        mixed += valid ? 1u : 0u;
    }
    wave_count(&V.c->mixed, mixed);
}

// Shadow ray i of the current bounce (scene record B): adds a visible
// contribution to the survivor's radiance in set 1 - p or to the ended
// path's entry.
template <bool kInterfaces>
RP_DEV void pool_shadow_ray(const PoolView &V, uint32_t p, Stack &st, uint32_t i, const BatchScene &B) {
    const WfState &N = V.S[1 - p];
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
    Wavelengths wl = path_wavelengths(B.S, meta.z);
    float tr[4];
    bool vis = transmittance<kInterfaces>(B.S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr, B.P.boundary);
    if (vis) {
        float4 *dst = (meta.w & kEnded) ? &V.F.L_end[meta.w & ~kEnded] : &N.L[meta.w];
        float4 L = *dst;
        PathState ps;
        ps.L[0] = L.x, ps.L[1] = L.y, ps.L[2] = L.z, ps.L[3] = L.w;
        shadow_apply(ps, sr, vis, tr);
        *dst = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
    }
}

// kInterfaces = false when no scene of the batch has null interfaces (the
// visibility test then needs no interface walk, and far fewer registers).
// One thread per shadow ray; the grid covers the capacity.
template <int BS, bool kInterfaces>
__global__ void __launch_bounds__(BS) k_pool_shadow(PoolView V, uint32_t p) {
    __shared__ uint32_t s_node[kShortStack * BS];
    __shared__ float s_t[kShortStack * BS];
    const uint32_t n = V.c->sn, nb = (n + BS - 1) / BS;
    if (blockIdx.x >= nb) return;
    RP_STACK(st);
    st.init(s_node, s_t, BS);
    const uint32_t i = blockIdx.x * BS + threadIdx.x;
    const bool valid = i < n;
    const uint32_t s = valid ? V.sh_sid[i] : 0u;
    bool mine;
    const uint32_t u = wave_scene(valid, s, mine);
    const bool other = valid && !mine;
    const uint32_t o = wave_append(other, &V.c->ovs);
    if (other) V.ov[o] = i;
    if (mine) pool_shadow_ray<kInterfaces>(V, p, st, i, uniform_scene(V.scenes, u));
}

// The shadow rays k_pool_shadow handed over, one scene per lane.
template <bool kInterfaces>
__global__ void __launch_bounds__(256) k_pool_shadow_mixed(PoolView V, uint32_t p) {
    __shared__ uint32_t s_node[kShortStack * 256];
    __shared__ float s_t[kShortStack * 256];
    RP_STACK(st);
    st.init(s_node, s_t, 256);
    const uint32_t n = V.c->ovs;
    uint32_t rays = 0;
    RP_GRID_STRIDE(n) {
        const uint32_t i = base_ + threadIdx.x;
        if (i < n) {
            const uint32_t j = V.ov[i];
            pool_shadow_ray<kInterfaces>(V, p, st, j, V.scenes[V.sh_sid[j]]);
            ++rays;
        }
    }
    wave_count(&V.c->mixed, rays);
}

// Ended paths: sensor XYZ into their scene's per-path planes.  All en ended
// paths retire here, and the bounce's shadow rays are final: one thread
// adds both counts.
__global__ void k_pool_retire(PoolView V) {
    const uint32_t n = V.c->en;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        V.c->done += n;
        V.c->rays[1] += V.c->sn;
    }
    uint32_t failed = 0;
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
        }
    }
    wave_count(&V.c->rays[2], failed);
}

// Latency tail of the pool: once every path has been issued and at most a
// tail's worth are live (state set p, all awaiting their next closest hit),
// one thread finishes each -- the events of the bounce kernels in the same
// order, then sensor XYZ into its scene's planes as k_pool_retire writes it.
// Ray and failure counts go to the pool counters; the host counts the n
// paths as done.
template <bool kInterfaces>
__global__ void __launch_bounds__(kWfBlock) RP_PATH_ATTR k_pool_tail(PoolView V, uint32_t p, uint32_t n) {
    __shared__ uint32_t s_node[kShortStack * kWfBlock];
    __shared__ float s_t[kShortStack * kWfBlock];
    RP_STACK(st);
    st.init(s_node, s_t, kWfBlock);
    const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
    uint32_t ext = 0, shr = 0, failed = 0;
    if (i < n) {
        const uint32_t sw = V.sid[p][i];
        for_scene(V.scenes, sw & ~kSidFresh, [&](const BatchScene &B) {
            PathState ps;
            if (sw & kSidFresh) load_compact(B.S, B.P, V.S[p], i, B.sample_begin, ps);
            else load_state(V.S[p], i, ps);
            const Wavelengths wl = path_wavelengths(B.S, ps.w2);
            run_path<kInterfaces>(B.S, B.P, st, ps, wl, ext, shr);
            float v[3] = {0, 0, 0};
            if (ps.status == RP_PATH_SUCCESS) sensor_xyz(B.S, wl, ps.L, v);
            else failed += 1;
            const uint64_t lane = (ps.sample - B.sample_begin) * B.npx + ps.pixel;
            B.xyz[lane] = v[0];
            B.xyz[B.paths + lane] = v[1];
            B.xyz[2 * B.paths + lane] = v[2];
        });
    }
    wave_count(&V.c->rays[0], ext);
    wave_count(&V.c->rays[1], shr);
    wave_count(&V.c->rays[2], failed);
}

} // namespace rpd
