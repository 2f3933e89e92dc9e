// Wavefront form of the path tracer: generate -> extend -> (sort) -> shade ->
// shadow -> accumulate, with path state in HBM (structure of arrays, 16-byte
// fields) compacted every bounce: two state sets, survivors appended at
// block-aggregated ballot positions (the frame path here and the batch pool
// in rp_batch.h).  The per-event arithmetic is rp_path.h's, so a wavefront frame
// equals the megakernel frame bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "rp_path.h"

namespace rpd {

// One set of path state (structure of arrays; passed by value to kernels).
struct WfState {
    float4 *o_t, *d, *beta, *density, *prev_density, *prev_pdf, *eta, *L, *prev_p;
    uint4 *id, *meta, *meta2;
};

// ------------------------------------------------------------ state I/O
RP_DEV void load_state(const WfState &B, uint32_t i, PathState &ps) {
    float4 a = B.o_t[i], b = B.d[i];
    ps.o = {a.x, a.y, a.z};
    ps.tmax = a.w;
    ps.d = {b.x, b.y, b.z};
    ps.prev_nc = __float_as_uint(b.w);
    float4 v;
    v = B.beta[i];
    ps.beta[0] = v.x, ps.beta[1] = v.y, ps.beta[2] = v.z, ps.beta[3] = v.w;
    v = B.density[i];
    ps.density[0] = v.x, ps.density[1] = v.y, ps.density[2] = v.z, ps.density[3] = v.w;
    v = B.prev_density[i];
    ps.prev_density[0] = v.x, ps.prev_density[1] = v.y, ps.prev_density[2] = v.z, ps.prev_density[3] = v.w;
    v = B.prev_pdf[i];
    ps.prev_pdf[0] = v.x, ps.prev_pdf[1] = v.y, ps.prev_pdf[2] = v.z, ps.prev_pdf[3] = v.w;
    v = B.eta[i];
    ps.eta_scale[0] = v.x, ps.eta_scale[1] = v.y, ps.eta_scale[2] = v.z, ps.eta_scale[3] = v.w;
    v = B.L[i];
    ps.L[0] = v.x, ps.L[1] = v.y, ps.L[2] = v.z, ps.L[3] = v.w;
    v = B.prev_p[i];
    ps.prev_p = {v.x, v.y, v.z};
    ps.primary = __float_as_uint(v.w);
    uint4 u = B.id[i];
    ps.pixel = u.x;
    ps.sample = uint64_t(u.y) | (uint64_t(u.z) << 32);
    ps.w2 = u.w;
    u = B.meta[i];
    ps.index = uint64_t(u.x) | (uint64_t(u.y) << 32);
    ps.count = u.z;
    ps.crossings = u.w;
    u = B.meta2[i];
    ps.medium = u.x;
    ps.mdraw = u.y;
    ps.flags = u.z;
    ps.status = u.w;
}
// kIds: also write the identity word (pixel, sample, wavelength word).  It
// never changes after generation, so only an in-place update may skip it;
// a path that moves (compaction) needs it at its new position.
template <bool kIds = true>
RP_DEV void store_state(const WfState &B, uint32_t i, const PathState &ps) {
    B.o_t[i] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.tmax);
    B.d[i] = make_float4(ps.d.x, ps.d.y, ps.d.z, __uint_as_float(ps.prev_nc));
    B.beta[i] = make_float4(ps.beta[0], ps.beta[1], ps.beta[2], ps.beta[3]);
    B.density[i] = make_float4(ps.density[0], ps.density[1], ps.density[2], ps.density[3]);
    B.prev_density[i] = make_float4(ps.prev_density[0], ps.prev_density[1], ps.prev_density[2], ps.prev_density[3]);
    B.prev_pdf[i] = make_float4(ps.prev_pdf[0], ps.prev_pdf[1], ps.prev_pdf[2], ps.prev_pdf[3]);
    B.eta[i] = make_float4(ps.eta_scale[0], ps.eta_scale[1], ps.eta_scale[2], ps.eta_scale[3]);
    B.L[i] = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
    B.prev_p[i] = make_float4(ps.prev_p.x, ps.prev_p.y, ps.prev_p.z, __uint_as_float(ps.primary));
    if (kIds) B.id[i] = make_uint4(ps.pixel, uint32_t(ps.sample), uint32_t(ps.sample >> 32), ps.w2);
    B.meta[i] = make_uint4(uint32_t(ps.index), uint32_t(ps.index >> 32), ps.count, ps.crossings);
    B.meta2[i] = make_uint4(ps.medium, ps.mdraw, ps.flags, ps.status);
}

// Wave-aggregated queue append: one atomic per wave that has a lane to
// append (for rare appends; the hot queues use block_append below).
RP_DEV uint32_t wave_append(bool pred, uint32_t *counter) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    uint32_t lane = threadIdx.x & 63;
    uint32_t leader = uint32_t(__builtin_ctzll(mask));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, uint32_t(__builtin_popcountll(mask)));
    base = __shfl(base, int(leader));
    uint64_t below = mask & ((uint64_t(1) << lane) - 1);
    return base + uint32_t(__builtin_popcountll(below));
}

RP_DEV void wave_count(unsigned long long *counter, uint32_t v) {
    unsigned long long e = v;
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off);
    if ((threadIdx.x & 63) == 0 && e) atomicAdd(counter, e);
}

// Atomics on one L2 cache line are serialised: measured on MI355X, one
// returning atomicAdd per wave on a single counter costs ~11 ns per wave
// (518,400 waves: 5.9 ms; three counters on one line: 17.6 ms; three
// counters on separate 256-B lines: 6.0 ms; one atomic per 256-lane block:
// 1.5 ms).  Hence: every statistics counter has its own 256-B line (counter
// k at counters[k * kCounterStride]), queue counters that are appended to
// in the same kernel sit on separate lines, and the appends are aggregated
// per block.
constexpr int kCounterStride = 32; // unsigned long long entries per line
constexpr int kQueueStride = 64;   // uint32_t entries per line

// LDS of block_append: per-wave counts (then per-wave offsets) and the
// block's base position per queue.
template <int BS, int K> struct AppendLds {
    uint32_t cnt[K][BS / 64];
    uint32_t base[K];
};

// Block-aggregated append to K queues: pos[k] is this thread's position in
// queue k when pred[k] holds (positions of one block are contiguous, in
// lane order).  One atomic per queue and block.  Every thread of the block
// must call it (it synchronises the block twice).
template <int BS, int K>
RP_DEV void block_append(AppendLds<BS, K> &sh, const bool (&pred)[K], uint32_t *const (&queue)[K],
                         uint32_t (&pos)[K]) {
    static_assert(BS % 64 == 0 && K <= 64, "block_append shape");
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    uint64_t m[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        m[k] = __ballot(pred[k]);
        if (lane == 0) sh.cnt[k][wave] = uint32_t(__builtin_popcountll(m[k]));
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (threadIdx.x == uint32_t(k)) {
            uint32_t acc = 0;
            for (int w = 0; w < BS / 64; ++w) {
                const uint32_t c = sh.cnt[k][w];
                sh.cnt[k][w] = acc;
                acc += c;
            }
            sh.base[k] = acc ? atomicAdd(queue[k], acc) : 0u;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) pos[k] = sh.base[k] + sh.cnt[k][wave] + uint32_t(__builtin_popcountll(m[k] & below));
}

constexpr int kWfBlock = 128;

// Blocks keep the hardware's round-robin spread over the 8 XCDs (block b
// on XCD b % 8).  Giving each XCD a contiguous range of queue positions, so
// that coherent rays share one L2, measured slower on the 1080p terrain frame
// (897.6 vs 927.5 M paths/s): neighbouring rays cost alike, and contiguous
// ranges leave some XCDs with the expensive part of the image.

// ------------------------------------------------------------ frame wavefront
// The frame path keeps the live paths' state physically compacted: bounce k
// reads state set C[0, qn) and the shading kernel appends the surviving paths
// to set N (block-aggregated positions), so every state, hit and shadow access
// is a dense, coalesced stream.  A path that ends leaves its radiance and
// (status, wavelength word) at its lane (pixel + npx * sample offset); a
// shadow ray knows where its contribution goes: the survivor's new position,
// or (high bit set) the lane of a path that ended at this event.
struct FrameWf {
    float4 *hit_t;
    uint2 *hit_id;
    float4 *sh_o, *sh_d, *sh_val;
    uint4 *sh_meta; // (medium, no-MIS flag, wavelength word, destination)
    float4 *L_end;  // per lane
    uint2 *end;     // per lane: (status, wavelength word)
    uint32_t *perm; // material-sorted order of [0, qn)
    uint32_t *qcount; // [0] survivors, [kQueueStride] shadow rays (separate lines)
    uint32_t *hist;   // 256 bucket counts, then 256 cursors; entry k at hist[k * kQueueStride]
    unsigned long long *tail; // latency tail's extension [0] and shadow [kCounterStride] rays
};
constexpr uint32_t kEnded = 0x80000000u;

// Pixel of position q (< W*H) within one sample of a frame: 8x8 pixel tiles
// in row-major tile order when both image sizes are multiples of 8, so that
// a wave's 64 paths start from one tile (coherent camera rays, nearby
// first hits), else row-major.  Results are stored per (sample, pixel), so
// the order changes nothing in the film.
RP_DEV uint32_t tile_pixel(const Scene &S, uint32_t q) {
    const uint32_t W = S.camera.width, H = S.camera.height;
    if ((W | H) & 7u) return q;
    const uint32_t t = q >> 6, l = q & 63u, tw = W >> 3;
    return ((t / tw) * 8u + (l >> 3)) * W + (t % tw) * 8u + (l & 7u);
}

// A new path's state, compactly: what path_init cannot rebuild (direction
// and wavelength word in C.d; Sobol index, pixel and sample offset from
// sample_begin in C.meta).
RP_DEV void store_compact(const WfState &C, uint32_t i, const PathState &ps, uint64_t sample_begin) {
    C.d[i] = make_float4(ps.d.x, ps.d.y, ps.d.z, __uint_as_float(ps.w2));
    C.meta[i] = make_uint4(uint32_t(ps.index), uint32_t(ps.index >> 32), ps.pixel, uint32_t(ps.sample - sample_begin));
}
// The state of the new path stored compactly at C[j]: equal to path_begin's.
RP_DEV void load_compact(const Scene &S, const PathParams &P, const WfState &C, uint32_t j, uint64_t sample_begin,
                         PathState &ps) {
    const float4 d = C.d[j];
    const uint4 m = C.meta[j];
    path_init(S, P, m.z, sample_begin + m.w, ps);
    ps.d = {d.x, d.y, d.z};
    ps.w2 = __float_as_uint(d.w);
    ps.index = uint64_t(m.x) | (uint64_t(m.y) << 32);
}

// Camera samples of lanes [0, n).  kCompact: only what path_init cannot
// rebuild -- direction and wavelength word (C.d), Sobol index (C.meta) --
// 32 B per path instead of the 192-B state; the first bounce's kernels
// (kGen) read those and rebuild the rest.
template <bool kCompact = false>
__global__ void __launch_bounds__(256) k_wf_generate(Scene S, PathParams P, uint32_t npx, uint64_t sample_begin,
                                                       uint64_t n, WfState C) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PathState ps;
    uint32_t w0, w1;
    path_begin(S, P, tile_pixel(S, uint32_t(i % npx)), sample_begin + i / npx, ps, w0, w1);
    if (kCompact) store_compact(C, uint32_t(i), ps, sample_begin);
    else store_state(C, uint32_t(i), ps);
}

// Closest hit of position i of C[0, qn).  kGen (the first bounce, compact
// generation): the ray leaves the camera origin with the direction in C.d.
// With kCount, node fetches, triangle tests and stack overflows are tallied
// in counters 3, 4 and 7 (the instrumented replay); ray counts are the queue
// counts, summed on the host.
template <bool kCount, bool kGen>
RP_DEV void extend_ray(const Scene &S, Stack &st, const WfState &C, const FrameWf &F, uint32_t i, uint32_t qn,
                       unsigned long long *counters) {
    uint32_t nodes = 0, tris = 0, spills = 0;
    if (i < qn) {
        float4 a, b = C.d[i];
        if (kGen) a = make_float4(S.camera.origin[0], S.camera.origin[1], S.camera.origin[2], __builtin_inff());
        else a = C.o_t[i];
        Hit h;
        TraversalCounts tc = {0, 0, 0};
        traverse<false, kCount>(S, st, F3{a.x, a.y, a.z}, F3{b.x, b.y, b.z}, a.w, h, &tc);
        F.hit_t[i] = make_float4(h.t, h.b0, h.b1, h.b2);
        F.hit_id[i] = make_uint2(h.prim, h.id);
        if (kCount) {
            nodes = tc.nodes;
            tris = tc.tris;
            spills = tc.spills;
        }
    }
    if (kCount) {
        wave_count(&counters[3 * kCounterStride], nodes);
        wave_count(&counters[4 * kCounterStride], tris);
        wave_count(&counters[7 * kCounterStride], spills);
    }
}

// Closest hit for C[0, qn).
template <bool kCount, bool kGen = false>
__global__ void __launch_bounds__(kWfBlock) k_wf_extend(Scene S, WfState C, FrameWf F, uint32_t qn,
                                                          unsigned long long *counters) {
    __shared__ uint32_t s_node[kShortStack * kWfBlock];
    __shared__ float s_t[kShortStack * kWfBlock];
    RP_STACK(st);
    st.init(s_node, s_t, kWfBlock);
    extend_ray<kCount, kGen>(S, st, C, F, blockIdx.x * kWfBlock + threadIdx.x, qn, counters);
}

RP_DEV uint32_t material_key(const Scene &S, const uint2 *hit_id, uint32_t slot, uint32_t bits) {
    uint32_t prim = hit_id[slot].x;
    if (prim == kEmpty) return 0;
    // one gather (the triangle's material) and the small material table,
    // instead of triangle -> object -> material, three dependent loads
    uint32_t m = S.tri_material[prim];
    uint32_t kind = S.materials[m].kind & 3u;
    uint32_t K = 1u << bits;
    if (bits <= 2) return 1 + kind < K ? 1 + kind : K - 1;
    uint32_t key = 1 + (kind << (bits - 2)) + (m & ((1u << (bits - 2)) - 1));
    return key < K ? key : K - 1;
}

// Counting sort of queue entries by material key; `q` = nullptr sorts the
// positions [0, qn) themselves.  Each thread takes kSortItems entries
// (coalesced: entry j of a thread at block_base + j * 256 + thread), so a
// block covers 4096 entries and adds one global atomic per key: atomics on
// one line serialise (~11 ns each), and at one block per 256 entries the
// two passes over a 33 M-entry bounce spent ~2.8 ms in them.
constexpr uint32_t kSortItems = 16;
constexpr uint32_t kSortBlockItems = 256 * kSortItems;
inline uint32_t sort_blocks(uint32_t n) { return (n + kSortBlockItems - 1) / kSortBlockItems; }
__global__ void __launch_bounds__(256) k_sort_hist(Scene S, const uint2 *hit_id, const uint32_t *q, uint32_t qn,
                                                     uint32_t bits, uint32_t *hist) {
    __shared__ uint32_t h[256];
    uint32_t K = 1u << bits;
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const uint32_t first = blockIdx.x * kSortBlockItems + threadIdx.x;
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t i = first + j * 256;
        if (i < qn) atomicAdd(&h[material_key(S, hit_id, q ? q[i] : i, bits)], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
        if (h[k]) atomicAdd(&hist[k * kQueueStride], h[k]);
}
__global__ void k_sort_scan(uint32_t bits, uint32_t *hist) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t K = 1u << bits, acc = 0;
    for (uint32_t k = 0; k < K; ++k) {
        uint32_t c = hist[k * kQueueStride];
        hist[(256 + k) * kQueueStride] = acc; // cursor
        acc += c;
    }
}
__global__ void __launch_bounds__(256) k_sort_scatter(Scene S, const uint2 *hit_id, const uint32_t *q, uint32_t qn,
                                                        uint32_t bits, uint32_t *hist, uint32_t *out) {
    __shared__ uint32_t h[256], base[256];
    uint32_t K = 1u << bits;
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) h[k] = 0;
    __syncthreads();
    const uint32_t first = blockIdx.x * kSortBlockItems + threadIdx.x;
    uint32_t packed[kSortItems]; // key << 12 | rank within the block (< 4096)
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t i = first + j * 256;
        packed[j] = 0;
        if (i < qn) {
            const uint32_t key = material_key(S, hit_id, q ? q[i] : i, bits);
            packed[j] = key << 12 | atomicAdd(&h[key], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x)
        base[k] = h[k] ? atomicAdd(&hist[(256 + k) * kQueueStride], h[k]) : 0;
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kSortItems; ++j) {
        const uint32_t i = first + j * 256;
        if (i < qn) out[base[packed[j] >> 12] + (packed[j] & 4095u)] = q ? q[i] : i;
    }
}
__global__ void k_wf_reset(uint32_t *qcount, uint32_t *hist, uint32_t bits) {
    uint32_t t = threadIdx.x;
    if (t < 2) qcount[t * kQueueStride] = 0;
    uint32_t K = 1u << bits;
    for (uint32_t k = t; k < K; k += blockDim.x) {
        hist[k * kQueueStride] = 0;
        hist[(256 + k) * kQueueStride] = 0;
    }
}

// Minimum waves per SIMD requested for the shading kernel (register budget
// 512 / waves); 0 leaves the choice to the compiler (171 VGPRs, 2 waves).
// Measured on the 1080p terrain frame (wavefront / sorted): compiler's
// choice 31.7 / 31.2 ms, 3 waves 29.9 / 30.6 ms, 4 waves 31.6 / 32.0 ms.
#ifndef RP_SHADE_WAVES
#define RP_SHADE_WAVES 3
#endif
// Lanes per shading block (one returning append atomic per queue and
// block).  Measured on the material-sorted 1080p frame: 128 / 256 / 512
// lanes 24.66 / 24.67 / 29.53 ms; the appends do not limit shading, and 512
// lanes leave room for only 2 waves per SIMD at 168 VGPRs.
#ifndef RP_SHADE_BLOCK
#define RP_SHADE_BLOCK 128
#endif
constexpr int kShadeBlock = RP_SHADE_BLOCK;
#if RP_SHADE_WAVES > 0
#define RP_SHADE_ATTR __attribute__((amdgpu_waves_per_eu(RP_SHADE_WAVES)))
#else
#define RP_SHADE_ATTR
#endif

// One path event for the path at C[j] when `valid`: survivors are appended
// to N and shadow rays to the shadow stream (every thread of the block must
// call this together: the appends are block-aggregated).  kGen (the first
// bounce): the path at j is lane j's, compactly generated.
template <bool kGen>
RP_DEV void shade_one(const Scene &S, const PathParams &P, const WfState &C, const WfState &N, const FrameWf &F,
                      bool valid, uint32_t j, uint32_t npx, uint64_t sample_begin, AppendLds<kShadeBlock, 2> &sh) {
    bool cont = false, shadow = false;
    PathState ps;
    ShadowReq sr;
    sr.active = false;
    uint32_t lane = 0;
    if (valid) {
        if (kGen) load_compact(S, P, C, j, sample_begin, ps);
        else load_state(C, j, ps);
        float4 ht = F.hit_t[j];
        uint2 hi = F.hit_id[j];
        Hit h;
        h.prim = hi.x;
        h.id = hi.y;
        h.t = ht.x;
        h.b0 = ht.y;
        h.b1 = ht.z;
        h.b2 = ht.w;
        NoHook nh;
        cont = path_event(S, P, ps, h, sr, nh);
        shadow = sr.active;
        cont = cont && path_alive(S, ps);
        lane = uint32_t(ps.sample - sample_begin) * npx + ps.pixel;
    }
    const bool pred[2] = {cont, shadow};
    uint32_t *const queue[2] = {&F.qcount[0], &F.qcount[kQueueStride]};
    uint32_t at[2];
    block_append(sh, pred, queue, at);
    const uint32_t pos = at[0], spos = at[1];
    if (cont) {
        store_state(N, pos, ps);
    } else if (valid) {
        F.L_end[lane] = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
        F.end[lane] = make_uint2(ps.status, ps.w2);
    }
    if (shadow) {
        F.sh_o[spos] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmax);
        F.sh_d[spos] = make_float4(sr.d.x, sr.d.y, sr.d.z, sr.w);
        F.sh_val[spos] = make_float4(sr.val[0], sr.val[1], sr.val[2], sr.val[3]);
        F.sh_meta[spos] = make_uint4(sr.medium, sr.delta_or_nomis ? 1u : 0u, ps.w2, cont ? pos : (kEnded | lane));
    }
}

// Any material: reads C[i], or C[perm[i]] in material-sorted order.
template <bool kGen = false>
__global__ void __launch_bounds__(kShadeBlock) RP_SHADE_ATTR k_wf_shade(Scene S, PathParams P, WfState C, WfState N,
                                                                     FrameWf F, uint32_t qn, bool sorted, uint32_t npx,
                                                                     uint64_t sample_begin) {
    __shared__ AppendLds<kShadeBlock, 2> sh;
    const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
    const bool valid = i < qn;
    shade_one<kGen>(S, P, C, N, F, valid, valid ? (sorted ? F.perm[i] : i) : 0u, npx, sample_begin, sh);
}

// Shadow ray i of one bounce's sn; adds a visible contribution to the
// survivor's radiance in N or to an ended path's lane.  With kCount
// (instrumented replay) node fetches, triangle tests and stack overflows are
// tallied in counters 5, 6 and 7; shadow ray counts are the queue counts,
// summed on the host.
template <bool kInterfaces, bool kCount>
RP_DEV void shadow_ray(const Scene &S, const PathParams &P, Stack &st, const WfState &N, const FrameWf &F, uint32_t i,
                       uint32_t sn, unsigned long long *counters) {
    TraversalCounts tc = {0, 0, 0};
    if (i < sn) {
        float4 o = F.sh_o[i], d = F.sh_d[i], v = F.sh_val[i];
        uint4 meta = F.sh_meta[i];
        ShadowReq sr;
        sr.o = {o.x, o.y, o.z};
        sr.tmax = o.w;
        sr.d = {d.x, d.y, d.z};
        sr.w = d.w;
        sr.val[0] = v.x, sr.val[1] = v.y, sr.val[2] = v.z, sr.val[3] = v.w;
        sr.medium = meta.x;
        sr.delta_or_nomis = meta.y != 0;
        Wavelengths wl = path_wavelengths(S, meta.z);
        float tr[4];
        bool vis = transmittance<kInterfaces, kCount>(S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr, P.boundary,
                                                      nullptr, RP_NONE, &tc);
        if (vis) {
            float4 *dst = (meta.w & kEnded) ? &F.L_end[meta.w & ~kEnded] : &N.L[meta.w];
            float4 L = *dst;
            PathState ps;
            ps.L[0] = L.x, ps.L[1] = L.y, ps.L[2] = L.z, ps.L[3] = L.w;
            shadow_apply(ps, sr, vis, tr);
            *dst = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
        }
    }
    if (kCount) {
        wave_count(&counters[5 * kCounterStride], tc.nodes);
        wave_count(&counters[6 * kCounterStride], tc.tris);
        wave_count(&counters[7 * kCounterStride], tc.spills);
    }
}

// Shadow rays of one bounce.
template <bool kInterfaces, bool kCount = false>
__global__ void __launch_bounds__(kWfBlock) k_wf_shadow(Scene S, PathParams P, WfState N, FrameWf F, uint32_t sn,
                                                          unsigned long long *counters) {
    __shared__ uint32_t s_node[kShortStack * kWfBlock];
    __shared__ float s_t[kShortStack * kWfBlock];
    RP_STACK(st);
    st.init(s_node, s_t, kWfBlock);
    shadow_ray<kInterfaces, kCount>(S, P, st, N, F, blockIdx.x * kWfBlock + threadIdx.x, sn, counters);
}

// One launch for a bounce's closest hits and the previous bounce's shadow
// rays (scenes without null interfaces): blocks [0, eb) trace the extension
// rays of C[0, qn), the rest the shadow rays [0, sn), which add to the
// radiance of C (the previous bounce's survivor set).  Neither half reads what
// the other writes (hit records against radiance and ended lanes).  The late
// bounces are latency-bound (a launch lasts as long as its slowest rays), and
// one launch instead of two overlaps those tails.
// Left to itself the compiler gives this kernel 80 VGPRs (6 waves per SIMD);
// measured on the 1080p terrain frame: 1.90 ms per launch, 1.84 ms at 7 waves
// (72 VGPRs), 1.85 ms at 8 waves (64 VGPRs).
template <bool kCount>
__global__ void __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(7)))
k_wf_trace(Scene S, PathParams P, WfState C, FrameWf F, uint32_t qn,
                                                         uint32_t eb, uint32_t sn, unsigned long long *counters) {
    __shared__ uint32_t s_node[kShortStack * kWfBlock];
    __shared__ float s_t[kShortStack * kWfBlock];
    RP_STACK(st);
    st.init(s_node, s_t, kWfBlock);
    if (blockIdx.x < eb) extend_ray<kCount, false>(S, st, C, F, blockIdx.x * kWfBlock + threadIdx.x, qn, counters);
    else shadow_ray<false, kCount>(S, P, st, C, F, (blockIdx.x - eb) * kWfBlock + threadIdx.x, sn, counters);
}

// ------------------------------------------------------------ latency tail
// Once few paths are live, a bounce costs its fixed overheads (four launches
// and the host round trip for the queue counts, ~50 us) more than its work.
// Then one thread finishes each remaining path of C[0, qn): closest hit,
// event, shadow ray, in the order the bounce kernels apply them (the shadow
// contribution is added after the event, as the shadow kernel adds it), so
// the radiance is unchanged bit for bit.
template <bool kInterfaces>
__global__ void __launch_bounds__(kWfBlock) RP_PATH_ATTR k_wf_tail(Scene S, PathParams P, WfState C, FrameWf F, uint32_t qn,
                                                       uint32_t npx, uint64_t sample_begin) {
    __shared__ uint32_t s_node[kShortStack * kWfBlock];
    __shared__ float s_t[kShortStack * kWfBlock];
    RP_STACK(st);
    st.init(s_node, s_t, kWfBlock);
    const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
    uint32_t ext = 0, shr = 0;
    if (i < qn) {
        PathState ps;
        load_state(C, i, ps);
        run_path<kInterfaces>(S, P, st, ps, path_wavelengths(S, ps.w2), ext, shr);
        const uint32_t lane = uint32_t(ps.sample - sample_begin) * npx + ps.pixel;
        F.L_end[lane] = make_float4(ps.L[0], ps.L[1], ps.L[2], ps.L[3]);
        F.end[lane] = make_uint2(ps.status, ps.w2);
    }
    wave_count(&F.tail[0], ext);
    wave_count(&F.tail[kCounterStride], shr);
}

__global__ void k_wf_zero_tail(unsigned long long *tail) {
    if (threadIdx.x < 2) tail[threadIdx.x * kCounterStride] = 0;
}

// Live paths at or below which the frame finishes in the latency tail.  By
// default: as many as the tail kernels hold resident at once (2 waves per
// SIMD at their ~230 VGPRs: CUs * 4 SIMDs * 2 * 64 lanes, 131072 on a
// 256-CU MI355X); measured on the 256x256 fog Cornell box, 16 spp, the
// inverse-rendering step took 10.00 / 9.04 / 8.51 / 9.13 ms with no tail /
// 65536 / 131072 / 262144, and the 512x512 volume frame 24.6 / 20.8 / 19.7 /
// 20.5 ms.
inline uint32_t tail_limit(const rp_frame_config &cfg) {
    // the instrumented replay counts every query in the bounce kernels
    if (cfg.flags & (RP_FRAME_NO_TAIL | RP_FRAME_COUNT_TRAFFIC)) return 0;
    if (cfg.tail_paths) return cfg.tail_paths;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 65536;
    return uint32_t(cus) * 4u * 2u * 64u;
}

__global__ void __launch_bounds__(256) k_wf_finish(Scene S, FrameWf F, uint64_t n, float *xyz,
                                                     unsigned long long *counters) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t failed = 0;
    if (i < n) {
        float4 L = F.L_end[i];
        uint2 e = F.end[i];
        uint32_t status = e.x;
        Wavelengths wl = path_wavelengths(S, e.y);
        float Lv[4] = {L.x, L.y, L.z, L.w};
        float v[3] = {0, 0, 0};
        if (status == RP_PATH_SUCCESS) sensor_xyz(S, wl, Lv, v);
        else failed = 1;
        xyz[i] = v[0];
        xyz[n + i] = v[1];
        xyz[2 * n + i] = v[2];
    }
    wave_count(&counters[2 * kCounterStride], failed);
}

#define WF_CHECK(expr)                                                                                        \
    do {                                                                                                       \
        hipError_t err_ = (expr);                                                                              \
        if (err_ != hipSuccess)                                                                                \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(err_));                     \
    } while (0)

// Two timing events owned by a scope: destroyed however the scope exits.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        WF_CHECK(hipEventCreate(&a));
        if (hipError_t e = hipEventCreate(&b); e != hipSuccess) {
            (void)hipEventDestroy(a);
            throw std::runtime_error(std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
};

template <typename T> inline void wf_alloc(T *&p, uint64_t n) { WF_CHECK(hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T))); }

// Owner of the frame wavefront's buffers: two compacted state sets, the
// per-position hit and shadow streams and the per-lane results.
struct FrameBuffers {
    WfState state[2]{}; // only the path-state fields are used
    FrameWf f{};
    uint64_t capacity = 0;
    uint32_t *host_count = nullptr; // pinned
    FrameBuffers() = default;
    FrameBuffers(const FrameBuffers &) = delete;
    FrameBuffers &operator=(const FrameBuffers &) = delete;
    ~FrameBuffers() { release(); }
    void release() {
        for (WfState &s : state) {
            void *ptrs[] = {s.o_t, s.d, s.beta, s.density, s.prev_density, s.prev_pdf, s.eta, s.L, s.prev_p,
                            s.id, s.meta, s.meta2};
            for (void *p : ptrs)
                if (p) (void)hipFree(p);
            s = WfState{};
        }
        void *ptrs[] = {f.hit_t, f.hit_id, f.sh_o, f.sh_d, f.sh_val, f.sh_meta, f.L_end, f.end, f.perm,
                        f.qcount, f.hist, f.tail};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        f = FrameWf{};
        if (host_count) (void)hipHostFree(host_count);
        host_count = nullptr;
        capacity = 0;
    }
    void ensure(uint64_t n);
    // The latency tail's extension and shadow rays since the last frame that
    // began with first_chunk (synchronous).
    void tail_rays(uint64_t out[2]) const {
        out[0] = out[1] = 0;
        if (!f.tail) return;
        unsigned long long v[kCounterStride + 1];
        WF_CHECK(hipMemcpy(v, f.tail, sizeof v, hipMemcpyDeviceToHost));
        out[0] = v[0];
        out[1] = v[kCounterStride];
    }
};

inline void FrameBuffers::ensure(uint64_t n) {
    if (n <= capacity) return;
    release();
    for (WfState &s : state) {
        wf_alloc(s.o_t, n), wf_alloc(s.d, n), wf_alloc(s.beta, n), wf_alloc(s.density, n);
        wf_alloc(s.prev_density, n), wf_alloc(s.prev_pdf, n), wf_alloc(s.eta, n), wf_alloc(s.L, n);
        wf_alloc(s.prev_p, n), wf_alloc(s.id, n), wf_alloc(s.meta, n), wf_alloc(s.meta2, n);
    }
    wf_alloc(f.hit_t, n), wf_alloc(f.hit_id, n);
    wf_alloc(f.sh_o, n), wf_alloc(f.sh_d, n), wf_alloc(f.sh_val, n), wf_alloc(f.sh_meta, n);
    wf_alloc(f.L_end, n), wf_alloc(f.end, n), wf_alloc(f.perm, n);
    wf_alloc(f.qcount, 2 * kQueueStride);
    wf_alloc(f.hist, 512 * kQueueStride); // every bucket counter on its own line
    wf_alloc(f.tail, 2 * kCounterStride);
    WF_CHECK(hipHostMalloc(reinterpret_cast<void **>(&host_count), 2 * kQueueStride * sizeof(uint32_t),
                           hipHostMallocDefault));
    capacity = n;
}

// Runs n paths (position i: pixel tile_pixel(i % npx), sample sample_begin +
// i / npx; npx = W * H) to
// completion and writes planar per-path XYZ.  Returns device seconds spent in
// the launches that trace extension rays (k_wf_extend, and k_wf_trace, which
// traces the previous bounce's shadow rays in the same launch); adds the
// extension and shadow ray counts of the bounce kernels (their queue counts)
// to rays[0] and rays[1]; the latency tail's go to the arena's tail
// counters, zeroed when first_chunk.  With `shadow`, the device seconds and
// launches of the shadow-only launches (k_wf_shadow) are added to shadow[0]
// and shadow[1]; with `fused`, the k_wf_trace launches and the shadow rays
// they traced to fused[0] and fused[1].
inline double wavefront_render(const Scene &S, const PathParams &P, FrameBuffers &FB, const rp_frame_config &cfg,
                               uint32_t npx, uint64_t sample_begin, uint64_t n, float *xyz,
                               unsigned long long *counters, hipStream_t stream, uint64_t &launches,
                               uint64_t rays[2], bool first_chunk, double *shadow = nullptr,
                               uint64_t *fused = nullptr) {
    if (n > (uint64_t(1) << 31)) throw std::runtime_error("wavefront chunk too large");
    FB.ensure(n);
    const FrameWf &F = FB.f;
    if (first_chunk) hipLaunchKernelGGL(k_wf_zero_tail, dim3(1), dim3(64), 0, stream, F.tail);
    const uint32_t tail = tail_limit(cfg);
    const bool sorted = cfg.mode == RP_MODE_WAVEFRONT_SORTED;
    // Default key width 4 (16 keys: miss, then kind and material bits).
    // Measured on the 1080p terrain frame: 1 / 2 / 3 / 4 bits 25.96 / 24.94 /
    // 24.81 / 24.71 ms, unsorted 27.4 ms.
    uint32_t bits = cfg.sort_bits ? cfg.sort_bits : 4;
    if (bits < 1 || bits > 8) throw std::runtime_error("sort_bits must be in [1, 8]");
    int cur = 0;
    // compact generation unless the frame starts in the tail (which loads
    // whole states)
    bool gen = n > tail;
    if (gen)
        hipLaunchKernelGGL(k_wf_generate<true>, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream, S, P, npx,
                           sample_begin, n, FB.state[cur]);
    else
        hipLaunchKernelGGL(k_wf_generate<false>, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream, S, P, npx,
                           sample_begin, n, FB.state[cur]);
    uint32_t qn = uint32_t(n);
    EventPair ev, ev_sh; // destroyed on every exit, exceptions included
    hipEvent_t e0 = ev.a, e1 = ev.b;
    const bool count = (cfg.flags & RP_FRAME_COUNT_TRAFFIC) != 0;
    // Shadow rays of the previous bounce, pending in F.sh_*: they add to the
    // radiance of the current set C.  Without null interfaces they ride in the
    // next bounce's closest-hit launch (k_wf_trace); otherwise, and before the
    // latency tail or once no path is left, they get a launch of their own.
    uint32_t pending = 0;
    double trace = 0, shadow_seconds = 0;
    uint64_t shadow_launches = 0;
    bool shadow_timed = false; // a shadow launch's events not read yet
    auto read_shadow_time = [&]() {
        if (!shadow_timed) return;
        WF_CHECK(hipEventSynchronize(ev_sh.b));
        float ms = 0;
        WF_CHECK(hipEventElapsedTime(&ms, ev_sh.a, ev_sh.b));
        shadow_seconds += ms * 1e-3;
        shadow_timed = false;
    };
    auto shadow_alone = [&](const WfState &dst, uint32_t sn) {
        read_shadow_time();
        const dim3 sg((sn + kWfBlock - 1) / kWfBlock);
        WF_CHECK(hipEventRecord(ev_sh.a, stream));
        if (S.has_interfaces && count)
            hipLaunchKernelGGL((k_wf_shadow<true, true>), sg, dim3(kWfBlock), 0, stream, S, P, dst, F, sn, counters);
        else if (S.has_interfaces)
            hipLaunchKernelGGL((k_wf_shadow<true, false>), sg, dim3(kWfBlock), 0, stream, S, P, dst, F, sn, counters);
        else if (count)
            hipLaunchKernelGGL((k_wf_shadow<false, true>), sg, dim3(kWfBlock), 0, stream, S, P, dst, F, sn, counters);
        else
            hipLaunchKernelGGL((k_wf_shadow<false, false>), sg, dim3(kWfBlock), 0, stream, S, P, dst, F, sn, counters);
        WF_CHECK(hipEventRecord(ev_sh.b, stream));
        shadow_timed = true;
        ++shadow_launches;
    };
    while (qn > 0) {
        const WfState &C = FB.state[cur], &N = FB.state[1 - cur];
        const dim3 grid((qn + kWfBlock - 1) / kWfBlock);
        if (qn <= tail) {
            if (pending) shadow_alone(C, pending);
            pending = 0;
            if (S.has_interfaces)
                hipLaunchKernelGGL(k_wf_tail<true>, grid, dim3(kWfBlock), 0, stream, S, P, C, F, qn, npx, sample_begin);
            else
                hipLaunchKernelGGL(k_wf_tail<false>, grid, dim3(kWfBlock), 0, stream, S, P, C, F, qn, npx, sample_begin);
            break;
        }
        if (pending && S.has_interfaces) {
            shadow_alone(C, pending);
            pending = 0;
        }
        WF_CHECK(hipEventRecord(e0, stream));
        if (pending) {
            const uint32_t eb = grid.x, sb = (pending + kWfBlock - 1) / kWfBlock;
            if (count)
                hipLaunchKernelGGL((k_wf_trace<true>), dim3(eb + sb), dim3(kWfBlock), 0, stream, S, P, C, F, qn, eb,
                                   pending, counters);
            else
                hipLaunchKernelGGL((k_wf_trace<false>), dim3(eb + sb), dim3(kWfBlock), 0, stream, S, P, C, F, qn, eb,
                                   pending, counters);
            if (fused) {
                fused[0] += 1;
                fused[1] += pending;
            }
            pending = 0;
        } else if (gen && count)
            hipLaunchKernelGGL((k_wf_extend<true, true>), grid, dim3(kWfBlock), 0, stream, S, C, F, qn, counters);
        else if (gen)
            hipLaunchKernelGGL((k_wf_extend<false, true>), grid, dim3(kWfBlock), 0, stream, S, C, F, qn, counters);
        else if (count)
            hipLaunchKernelGGL((k_wf_extend<true, false>), grid, dim3(kWfBlock), 0, stream, S, C, F, qn, counters);
        else
            hipLaunchKernelGGL((k_wf_extend<false, false>), grid, dim3(kWfBlock), 0, stream, S, C, F, qn, counters);
        WF_CHECK(hipEventRecord(e1, stream));
        ++launches;
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(256), 0, stream, F.qcount, F.hist, bits);
        // The first bounce stays in generation order: its waves are 8x8
        // pixel tiles whose camera rays already hit alike surfaces, and
        // sorting them cost 0.29 ms in passes for 0.09 ms of shading.
        const bool sort_now = sorted && !gen;
        if (sort_now) {
            const dim3 g(sort_blocks(qn));
            hipLaunchKernelGGL(k_sort_hist, g, dim3(256), 0, stream, S, F.hit_id, nullptr, qn, bits, F.hist);
            hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(64), 0, stream, bits, F.hist);
            hipLaunchKernelGGL(k_sort_scatter, g, dim3(256), 0, stream, S, F.hit_id, nullptr, qn, bits, F.hist, F.perm);
        }
        const dim3 sgrid((qn + kShadeBlock - 1) / kShadeBlock);
        if (gen)
            hipLaunchKernelGGL(k_wf_shade<true>, sgrid, dim3(kShadeBlock), 0, stream, S, P, C, N, F, qn, sort_now, npx,
                               sample_begin);
        else
            hipLaunchKernelGGL(k_wf_shade<false>, sgrid, dim3(kShadeBlock), 0, stream, S, P, C, N, F, qn, sort_now, npx,
                               sample_begin);
        gen = false;
        WF_CHECK(hipMemcpyAsync(FB.host_count, F.qcount, (kQueueStride + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                stream));
        WF_CHECK(hipStreamSynchronize(stream));
        float ms = 0;
        WF_CHECK(hipEventElapsedTime(&ms, e0, e1));
        trace += ms * 1e-3;
        read_shadow_time(); // a shadow launch of this bounce has finished with the stream
        const uint32_t next = FB.host_count[0], sn = FB.host_count[kQueueStride];
        rays[0] += qn;
        rays[1] += sn;
        if (next > qn || sn > qn) {
            throw std::runtime_error("wavefront queue counts out of range");
        }
        pending = sn;
        cur = 1 - cur;
        qn = next;
    }
    if (pending) shadow_alone(FB.state[cur], pending); // no path left to trace: the last shadow rays alone
    read_shadow_time();
    if (shadow) {
        shadow[0] += shadow_seconds;
        shadow[1] += double(shadow_launches);
    }
    hipLaunchKernelGGL(k_wf_finish, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, stream, S, F, n, xyz, counters);
    WF_CHECK(hipGetLastError());
    return trace;
}

} // namespace rpd
