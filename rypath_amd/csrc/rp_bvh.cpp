// Host SAH BVH builder: binned SAH binary tree, collapsed to a 4-wide tree of
// 64-byte quantised nodes (rp::QNode).  Replaces the reference's Metal
// acceleration-structure build (src/metal_geometry.mm:125-290), since CDNA4
// has no ray-tracing units and traversal runs in software.
#include "rp_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <stdexcept>
#include <thread>

namespace rp {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box &b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const float *p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    double area() const {
        double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
        if (!(dx >= 0)) return 0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Prim {
    Box box;
    float c[3];
    uint32_t id;
};

struct BNode {
    Box box;
    uint32_t left = 0, right = 0; // children (internal)
    uint32_t first = 0, count = 0; // leaf range (count > 0)
};

struct Builder {
    std::vector<Prim> prims;
    uint32_t leaf_size;
    uint32_t bins;

    // Builds the subtree over prims[begin, end) into `out` (local indices); returns the root.
    uint32_t build(std::vector<BNode> &out, uint32_t begin, uint32_t end) {
        struct Item { uint32_t node, begin, end; };
        std::vector<Item> stack;
        uint32_t root = uint32_t(out.size());
        out.push_back({});
        stack.push_back({root, begin, end});
        std::vector<Box> bin_box(3 * bins);
        std::vector<uint32_t> bin_count(3 * bins);
        std::vector<double> right_area(bins);
        std::vector<uint32_t> right_count(bins);
        while (!stack.empty()) {
            Item it = stack.back();
            stack.pop_back();
            Box box, cbox;
            for (uint32_t i = it.begin; i < it.end; ++i) {
                box.grow(prims[i].box);
                cbox.grow(prims[i].c);
            }
            BNode nd;
            nd.box = box;
            uint32_t n = it.end - it.begin;
            if (n <= leaf_size) {
                nd.first = it.begin;
                nd.count = n;
                out[it.node] = nd;
                continue;
            }
            // binned SAH over the three axes
            int best_axis = -1;
            uint32_t best_split = 0;
            double best_cost = INFINITY;
            std::fill(bin_count.begin(), bin_count.end(), 0u);
            std::fill(bin_box.begin(), bin_box.end(), Box{});
            float scale[3];
            for (int a = 0; a < 3; ++a) {
                float ext = cbox.hi[a] - cbox.lo[a];
                scale[a] = ext > 0 ? float(bins) / ext : 0.0f;
            }
            for (uint32_t i = it.begin; i < it.end; ++i)
                for (int a = 0; a < 3; ++a) {
                    if (scale[a] == 0.0f) continue;
                    uint32_t b = std::min(bins - 1, uint32_t((prims[i].c[a] - cbox.lo[a]) * scale[a]));
                    bin_count[a * bins + b]++;
                    bin_box[a * bins + b].grow(prims[i].box);
                }
            for (int a = 0; a < 3; ++a) {
                if (scale[a] == 0.0f) continue;
                Box acc;
                uint32_t cnt = 0;
                for (uint32_t b = bins - 1; b >= 1; --b) {
                    acc.grow(bin_box[a * bins + b]);
                    cnt += bin_count[a * bins + b];
                    right_area[b] = acc.area();
                    right_count[b] = cnt;
                }
                acc = Box{};
                cnt = 0;
                for (uint32_t b = 0; b + 1 < bins; ++b) {
                    acc.grow(bin_box[a * bins + b]);
                    cnt += bin_count[a * bins + b];
                    if (cnt == 0 || right_count[b + 1] == 0) continue;
                    double cost = acc.area() * cnt + right_area[b + 1] * right_count[b + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_split = b + 1;
                    }
                }
            }
            uint32_t mid;
            if (best_axis < 0) {
                mid = it.begin + n / 2; // degenerate centroids: split by index
            } else {
                int a = best_axis;
                float lo = cbox.lo[a], sc = scale[a];
                uint32_t split = best_split, nb = bins;
                auto pm = std::partition(prims.begin() + it.begin, prims.begin() + it.end, [&](const Prim &p) {
                    return std::min(nb - 1, uint32_t((p.c[a] - lo) * sc)) < split;
                });
                mid = uint32_t(pm - prims.begin());
                if (mid == it.begin || mid == it.end) mid = it.begin + n / 2;
            }
            uint32_t l = uint32_t(out.size());
            out.push_back({});
            out.push_back({});
            nd.left = l;
            nd.right = l + 1;
            out[it.node] = nd;
            stack.push_back({l, it.begin, mid});
            stack.push_back({l + 1, mid, it.end});
        }
        return root;
    }
};

// 2^(e-127) for an exponent byte e in [1, 254].
float pow2_byte(uint8_t e) {
    uint32_t bits = uint32_t(e) << 23;
    float s;
    std::memcpy(&s, &bits, 4);
    return s;
}

struct Collapser {
    const std::vector<BNode> &bn;
    std::vector<QNode> nodes;
    std::vector<uint32_t> depth_of;
    uint32_t max_depth = 0, max_stack = 0;
    uint64_t leaves = 0;

    explicit Collapser(const std::vector<BNode> &b) : bn(b) {}

    void quantize(QNode &q, const Box *boxes, int n) {
        Box u;
        for (int i = 0; i < n; ++i) u.grow(boxes[i]);
        for (int a = 0; a < 3; ++a) {
            float lo = u.lo[a], hi = u.hi[a];
            float mag = std::max(std::fabs(lo), std::fabs(hi));
            float pad = mag * 0x1p-20f + 1e-30f;
            lo -= pad;
            hi += pad;
            q.origin[a] = lo;
            float ext = hi - lo;
            int e2;
            std::frexp(std::max(ext / 255.0f, 1e-30f), &e2); // ext/255 <= 2^e2
            int e = std::min(std::max(e2 + 127, 1), 254);
            for (;;) {
                float s = pow2_byte(uint8_t(e));
                bool ok = true;
                for (int i = 0; i < n && ok; ++i) {
                    float cl = boxes[i].lo[a] - (std::fabs(boxes[i].lo[a]) * 0x1p-20f + 1e-30f);
                    float ch = boxes[i].hi[a] + (std::fabs(boxes[i].hi[a]) * 0x1p-20f + 1e-30f);
                    double fl = std::floor((double(cl) - double(lo)) / double(s));
                    double fh = std::ceil((double(ch) - double(lo)) / double(s));
                    int ql = int(std::max(0.0, std::min(255.0, fl)));
                    int qh = int(std::max(0.0, std::min(255.0, fh)));
                    while (ql > 0 && decode_lo(lo, uint8_t(e), uint8_t(ql)) > cl) --ql;
                    while (qh < 255 && decode_lo(lo, uint8_t(e), uint8_t(qh)) < ch) ++qh;
                    if (decode_lo(lo, uint8_t(e), uint8_t(ql)) > cl || decode_lo(lo, uint8_t(e), uint8_t(qh)) < ch) {
                        ok = false;
                        break;
                    }
                    q.qlo[a][i] = uint8_t(ql);
                    q.qhi[a][i] = uint8_t(qh);
                }
                if (ok) break;
                if (++e > 254) throw std::runtime_error("BVH quantisation overflow");
            }
            q.exponent[a] = uint8_t(e);
        }
    }

    // Emits the 4-wide node for binary node b; returns its index.
    uint32_t emit(uint32_t b, uint32_t depth, uint32_t stack_so_far) {
        std::vector<uint32_t> kids = {bn[b].left, bn[b].right};
        for (;;) {
            if (kids.size() >= 4) break;
            int best = -1;
            double best_area = -1;
            for (size_t i = 0; i < kids.size(); ++i)
                if (bn[kids[i]].count == 0 && bn[kids[i]].box.area() > best_area) {
                    best_area = bn[kids[i]].box.area();
                    best = int(i);
                }
            if (best < 0) break;
            uint32_t k = kids[best];
            kids.erase(kids.begin() + best);
            kids.push_back(bn[k].left);
            kids.push_back(bn[k].right);
        }
        uint32_t idx = uint32_t(nodes.size());
        nodes.push_back({});
        QNode q{};
        for (int i = 0; i < 4; ++i) {
            q.child[i] = kEmptyChild;
            q.count[i] = 0;
            for (int a = 0; a < 3; ++a) {
                q.qlo[a][i] = 255;
                q.qhi[a][i] = 0;
            }
        }
        Box boxes[4];
        int n = int(kids.size());
        for (int i = 0; i < n; ++i) boxes[i] = bn[kids[i]].box;
        quantize(q, boxes, n);
        q.child_count = uint8_t(n);
        max_depth = std::max(max_depth, depth);
        uint32_t st = stack_so_far + uint32_t(n - 1);
        max_stack = std::max(max_stack, st + 1);
        for (int i = 0; i < n; ++i) {
            const BNode &c = bn[kids[i]];
            if (c.count) {
                if (c.count > 15 || c.first > kLeafFirstMask) throw std::runtime_error("BVH leaf too large");
                q.child[i] = kLeafBit | (c.count << kLeafCountShift) | c.first;
                q.count[i] = uint8_t(c.count);
                ++leaves;
            }
        }
        nodes[idx] = q;
        for (int i = 0; i < n; ++i) {
            if (bn[kids[i]].count) continue;
            uint32_t ci = emit(kids[i], depth + 1, st);
            nodes[idx].child[i] = ci;
        }
        return idx;
    }
};

} // namespace

Bvh4 build_bvh4(const float *positions, const uint32_t *indices, uint32_t triangle_count, uint32_t leaf_size,
                uint32_t bins, uint32_t threads) {
    auto t0 = std::chrono::steady_clock::now();
    Bvh4 out;
    (void)threads;
    if (triangle_count == 0) return out;
    Builder b;
    // Default 2 triangles per leaf: measured wavefront frames (MI355X) with
    // leaves of 4 / 2 / 1: 1080p terrain 27.80 / 27.22 / 29.50 ms, volume
    // 512^2 32 spp 16.55 / 16.05 / 16.52 ms (64 bins), Cornell 512^2 64 spp
    // 12.45 / 11.53 ms.
    b.leaf_size = std::min(std::max(leaf_size ? leaf_size : 2u, 1u), 15u);
    b.bins = std::min(std::max(bins ? bins : 32u, 4u), 256u);
    b.prims.resize(triangle_count);
    for (uint32_t t = 0; t < triangle_count; ++t) {
        Prim &p = b.prims[t];
        for (int k = 0; k < 3; ++k) p.box.grow(positions + 3 * size_t(indices[3 * t + k]));
        for (int a = 0; a < 3; ++a) p.c[a] = 0.5f * (p.box.lo[a] + p.box.hi[a]);
        p.id = t;
    }
    std::vector<BNode> bn;
    bn.reserve(2 * size_t(triangle_count) / b.leaf_size + 16);
    uint32_t root = b.build(bn, 0, triangle_count);
    out.tri_order.resize(triangle_count);
    for (uint32_t i = 0; i < triangle_count; ++i) out.tri_order[i] = b.prims[i].id;
    Collapser c(bn);
    if (bn[root].count) {
        // Single leaf: wrap it in one 4-wide node.
        QNode q{};
        for (int i = 0; i < 4; ++i) {
            q.child[i] = kEmptyChild;
            for (int a = 0; a < 3; ++a) {
                q.qlo[a][i] = 255;
                q.qhi[a][i] = 0;
            }
        }
        Box boxes[1] = {bn[root].box};
        c.quantize(q, boxes, 1);
        q.child_count = 1;
        q.child[0] = kLeafBit | (bn[root].count << kLeafCountShift) | bn[root].first;
        q.count[0] = uint8_t(bn[root].count);
        c.nodes.push_back(q);
        c.max_depth = 1;
        c.max_stack = 1;
        c.leaves = 1;
    } else {
        c.nodes.reserve(bn.size() / 2 + 4);
        c.emit(root, 1, 0);
    }
    out.nodes = std::move(c.nodes);
    out.depth = c.max_depth;
    out.max_stack = c.max_stack;
    out.leaf_count = c.leaves;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

} // namespace rp
