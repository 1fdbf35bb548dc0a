// bvh_build.cpp -- binned-SAH binary BVH, collapsed to a W-wide BVH.
//
// The reference builds one binary BVH per mesh with a box-volume split rule,
// 16-triangle leaves and splits on multiples of 4 (src/core/primitive/mesh.cu:35-143),
// and loops linearly over meshes (src/core/meta/scene.hpp:39-46). Closest hit
// does not depend on the tree, so this build uses one SAH tree over all
// triangles of the scene, with small leaves and wide nodes sized for 64-lane
// wavefronts and 128-byte cache lines.
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>

#include "kr_layout.h"

namespace kr {
namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::numeric_limits<float>::infinity();
            hi[a] = -std::numeric_limits<float>::infinity();
        }
    }
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
    float area() const {
        float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
        if (!(e0 >= 0) || !(e1 >= 0) || !(e2 >= 0)) return 0.f;
        return 2.f * (e0 * e1 + e1 * e2 + e2 * e0);
    }
};

struct BinNode {
    Box box;
    uint32_t left = 0, right = 0;  // children (inner)
    uint32_t first = 0, count = 0; // leaf range into idx
    bool leaf = false;
};

struct Builder {
    const float *tv;
    uint32_t leaf_size;
    std::vector<Box> tbox;
    std::vector<float> cent;  // 3 per triangle
    std::vector<uint32_t> idx;
    std::vector<BinNode> nodes;

    static constexpr int kBins = 32;
    // Below this depth splits are binned SAH; deeper ones split at the object
    // median of the widest centroid axis, so that adversarial geometry (SAH
    // peeling one triangle per level) cannot make the tree, and this
    // recursion, O(n) deep: the binary depth stays below
    // kMaxSahDepth + log2(n) + 1.
    static constexpr uint32_t kMaxSahDepth = 40;

    int bin_of(uint32_t t, int ax, float lo, float scale) const {
        int b = int((cent[3 * size_t(t) + ax] - lo) * scale);
        return std::min(std::max(b, 0), kBins - 1);
    }

    uint32_t build(uint32_t begin, uint32_t end, uint32_t depth = 0) {
        uint32_t me = uint32_t(nodes.size());
        nodes.emplace_back();
        Box box, cb;
        box.reset();
        cb.reset();
        for (uint32_t i = begin; i < end; ++i) {
            box.grow(tbox[idx[i]]);
            cb.grow(&cent[3 * size_t(idx[i])]);
        }
        nodes[me].box = box;
        uint32_t n = end - begin;
        auto make_leaf_node = [&]() {
            nodes[me].leaf = true;
            nodes[me].first = begin;
            nodes[me].count = n;
            return me;
        };
        if (n <= 1) return make_leaf_node();
        int axis = 0;
        float ext[3] = {cb.hi[0] - cb.lo[0], cb.hi[1] - cb.lo[1], cb.hi[2] - cb.lo[2]};
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        uint32_t mid = begin + n / 2;
        bool split_found = false;
        if (depth >= kMaxSahDepth) {
            if (n <= leaf_size) return make_leaf_node();
        } else if (ext[axis] > 0.f) {
            // binned SAH over all three centroid axes (on the 1M-triangle
            // frame 5-9 % fewer node visits per ray than the widest axis alone)
            float best = std::numeric_limits<float>::infinity();
            int best_b = -1, best_axis = axis;
            for (int ax = 0; ax < 3; ++ax) {
                if (!(ext[ax] > 0.f)) continue;
                Box bins[kBins];
                uint32_t cnt[kBins] = {0};
                for (auto &b : bins) b.reset();
                const float scale = float(kBins) / ext[ax];
                for (uint32_t i = begin; i < end; ++i) {
                    int b = bin_of(idx[i], ax, cb.lo[ax], scale);
                    cnt[b]++;
                    bins[b].grow(tbox[idx[i]]);
                }
                float rarea[kBins];
                uint32_t rcnt[kBins];
                Box acc;
                acc.reset();
                uint32_t c = 0;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.grow(bins[b]);
                    c += cnt[b];
                    rarea[b] = acc.area();
                    rcnt[b] = c;
                }
                acc.reset();
                c = 0;
                for (int b = 0; b < kBins - 1; ++b) {
                    acc.grow(bins[b]);
                    c += cnt[b];
                    if (c == 0 || rcnt[b + 1] == 0) continue;
                    float cost = acc.area() * float(c) + rarea[b + 1] * float(rcnt[b + 1]);
                    if (cost < best) {
                        best = cost;
                        best_b = b;
                        best_axis = ax;
                    }
                }
            }
            float parea = box.area();
            float split_cost = 1.0f + (parea > 0 ? best / parea : float(n));
            float leaf_cost = float(n);
            if (n <= leaf_size && (best_b < 0 || leaf_cost <= split_cost)) return make_leaf_node();
            if (best_b >= 0) {
                const int ax = best_axis;
                const float lo = cb.lo[ax], scale = float(kBins) / ext[ax];
                auto it = std::partition(idx.begin() + begin, idx.begin() + end,
                                         [&](uint32_t t) { return bin_of(t, ax, lo, scale) <= best_b; });
                mid = uint32_t(it - idx.begin());
                split_found = mid > begin && mid < end;
            }
        } else if (n <= leaf_size) {
            return make_leaf_node();
        }
        if (!split_found) {
            // All centroids coincide on the axis (or binning failed): median by index.
            mid = begin + n / 2;
            if (ext[axis] > 0.f)
                std::nth_element(idx.begin() + begin, idx.begin() + mid, idx.begin() + end, [&](uint32_t a, uint32_t b) {
                    return cent[3 * size_t(a) + axis] < cent[3 * size_t(b) + axis];
                });
        }
        uint32_t l = build(begin, mid, depth + 1);
        uint32_t r = build(mid, end, depth + 1);
        nodes[me].left = l;
        nodes[me].right = r;
        return me;
    }
};

struct Collapser {
    const std::vector<BinNode> &bn;
    uint32_t W, NF;
    WideBvh &out;

    uint32_t emit(const std::vector<uint32_t> &kids, uint32_t depth) {
        uint32_t me = out.num_nodes++;
        out.nodes.resize(size_t(out.num_nodes) * NF, 0.f);
        if (depth > out.depth) out.depth = depth;
        uint32_t refs[8];
        for (uint32_t c = 0; c < W; ++c) {
            float *nd = &out.nodes[size_t(me) * NF];
            if (c >= kids.size()) {
                nd[0 * W + c] = nd[2 * W + c] = nd[4 * W + c] = std::numeric_limits<float>::infinity();
                nd[1 * W + c] = nd[3 * W + c] = nd[5 * W + c] = -std::numeric_limits<float>::infinity();
                refs[c] = kEmpty;
                continue;
            }
            const BinNode &k = bn[kids[c]];
            for (int a = 0; a < 3; ++a) {
                nd[(2 * a) * W + c] = k.box.lo[a];
                nd[(2 * a + 1) * W + c] = k.box.hi[a];
            }
            if (k.leaf) {
                refs[c] = make_leaf(k.first, k.count);
                out.num_leaves++;
            } else {
                refs[c] = 0;  // patched below
            }
        }
        for (uint32_t c = 0; c < kids.size(); ++c) {
            const BinNode &k = bn[kids[c]];
            if (!k.leaf) refs[c] = expand(kids[c], depth + 1);
        }
        float *nd = &out.nodes[size_t(me) * NF];
        std::memcpy(nd + 6 * W, refs, sizeof(uint32_t) * W);
        return me;
    }

    uint32_t expand(uint32_t b, uint32_t depth) {
        std::vector<uint32_t> kids = {bn[b].left, bn[b].right};
        while (kids.size() < W) {
            int pick = -1;
            float best = -1.f;
            for (size_t i = 0; i < kids.size(); ++i) {
                const BinNode &k = bn[kids[i]];
                if (k.leaf) continue;
                float a = k.box.area();
                if (a > best) {
                    best = a;
                    pick = int(i);
                }
            }
            if (pick < 0) break;
            uint32_t k = kids[pick];
            kids[pick] = bn[k].left;
            kids.push_back(bn[k].right);
        }
        return emit(kids, depth);
    }
};

}  // namespace

bool build_wide_bvh(const float *tri_verts, uint32_t num_tris, uint32_t width, uint32_t leaf_size, WideBvh &out,
                    std::string &err) {
    if (width != 2 && width != 4 && width != 8) {
        err = "bvh width must be 2, 4 or 8";
        return false;
    }
    // Default: one triangle per leaf. The walk's leaf step then tests exactly
    // one triangle in every lane, where leaves of up to 4 made a wave wait for
    // its largest leaf; on the 1M-triangle frame 3 % faster (100k-triangle
    // frame 1.5 %) for 1.7x the nodes.
    if (leaf_size == 0) leaf_size = 1;
    if (leaf_size > kMaxLeafTris) {
        err = "leaf size must be at most 16";
        return false;
    }
    if (num_tris == 0 || num_tris > kLeafFirstMask) {
        err = "triangle count out of range";
        return false;
    }
    Builder b;
    b.tv = tri_verts;
    b.leaf_size = leaf_size;
    b.tbox.resize(num_tris);
    b.cent.resize(3 * size_t(num_tris));
    b.idx.resize(num_tris);
    for (uint32_t t = 0; t < num_tris; ++t) {
        Box bx;
        bx.reset();
        for (int k = 0; k < 3; ++k) bx.grow(tri_verts + 9 * size_t(t) + 3 * k);
        b.tbox[t] = bx;
        for (int a = 0; a < 3; ++a) b.cent[3 * size_t(t) + a] = 0.5f * (bx.lo[a] + bx.hi[a]);
        b.idx[t] = t;
    }
    b.nodes.reserve(2 * size_t(num_tris) / std::max(1u, leaf_size) + 8);
    uint32_t root = b.build(0, num_tris);
    out = WideBvh();
    out.width = width;
    Collapser c{b.nodes, width, node_floats(width), out};
    if (b.nodes[root].leaf) {
        c.emit({root}, 1);
    } else {
        c.expand(root, 1);
    }
    out.tri_order = b.idx;
    return true;
}


namespace {
float qscale(int e) { return std::ldexp(1.0f, e); }
float qdecode(uint32_t q, float s, float p) { return std::fmaf(float(q), s, p); }
}  // namespace

void compress_wide_nodes(const WideBvh &bvh, std::vector<uint32_t> &out) {
    const uint32_t W = bvh.width, NF = node_floats(W), NW = qnode_bytes(W) / 4;
    out.assign(size_t(bvh.num_nodes) * NW, 0u);
    for (uint32_t n = 0; n < bvh.num_nodes; ++n) {
        const float *f = bvh.nodes.data() + size_t(n) * NF;
        uint32_t *o = out.data() + size_t(n) * NW;
        const uint32_t *ref = reinterpret_cast<const uint32_t *>(f + 6 * W);
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::numeric_limits<float>::infinity();
            hi[a] = -std::numeric_limits<float>::infinity();
        }
        for (uint32_t c = 0; c < W; ++c) {
            if (ref[c] == kEmpty) continue;
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], f[(2 * a) * W + c]);
                hi[a] = std::max(hi[a], f[(2 * a + 1) * W + c]);
            }
        }
        int e[3];
        float sc[3];
        for (int a = 0; a < 3; ++a) {
            if (!(lo[a] <= hi[a])) lo[a] = hi[a] = 0.f;  // no children (not produced by the builder)
            // 254 quanta cover the extent, leaving one for the round-up
            double ext = double(hi[a]) - double(lo[a]);
            int ea = ext > 0 ? int(std::ceil(std::log2(ext / 254.0))) : -125;
            // a quantum finer than one ulp of the coordinates resolves
            // nothing: keep the grid at least that coarse
            float m = std::max(std::fabs(lo[a]), std::fabs(hi[a]));
            float ulp = std::nextafter(m, std::numeric_limits<float>::infinity()) - m;
            if (ulp > 0) ea = std::max(ea, int(std::ceil(std::log2(double(ulp)))));
            ea = std::max(-125, std::min(125, ea));
            while (ea < 125 && double(qscale(ea)) * 254.0 < ext) ++ea;
            e[a] = ea;
            sc[a] = qscale(ea);
            std::memcpy(&o[a], &lo[a], 4);
        }
        o[3] = uint32_t(uint8_t(int8_t(e[0]))) | uint32_t(uint8_t(int8_t(e[1]))) << 8 |
               uint32_t(uint8_t(int8_t(e[2]))) << 16;
        if (W == 4) {
            // the 8 bytes of padding hold the scales themselves, which the
            // kernels read instead of building 2^e from the exponent bytes:
            // word 14 = the upper halves of s_x and s_y (a power of two is
            // exact in them), word 15 = s_z
            uint32_t sb[3];
            std::memcpy(sb, sc, sizeof sb);
            o[14] = (sb[0] >> 16) | (sb[1] & 0xFFFF0000u);
            o[15] = sb[2];
        }
        for (uint32_t c = 0; c < W; ++c) o[4 + c] = ref[c];
        uint8_t *q = reinterpret_cast<uint8_t *>(o + 4 + W);
        for (uint32_t c = 0; c < W; ++c) {
            if (ref[c] == kEmpty) continue;
            for (int a = 0; a < 3; ++a) {
                float cl = f[(2 * a) * W + c], ch = f[(2 * a + 1) * W + c];
                double rl = std::floor((double(cl) - double(lo[a])) / double(sc[a]));
                double rh = std::ceil((double(ch) - double(lo[a])) / double(sc[a]));
                int ql = int(std::max(0.0, std::min(255.0, rl)));
                int qh = int(std::max(0.0, std::min(255.0, rh)));
                // q s is exact and fma rounds once, monotonically, so the
                // floor / ceil above already contain the box; the loops only
                // guard the clamped ends
                while (ql > 0 && qdecode(uint32_t(ql), sc[a], lo[a]) > cl) --ql;
                while (qh < 255 && qdecode(uint32_t(qh), sc[a], lo[a]) < ch) ++qh;
                q[(2 * a) * W + c] = uint8_t(ql);
                q[(2 * a + 1) * W + c] = uint8_t(qh);
            }
        }
    }
}

void decode_qnode_box(const uint32_t *node, uint32_t W, uint32_t child, float lo[3], float hi[3]) {
    const uint8_t *q = reinterpret_cast<const uint8_t *>(node + 4 + W);
    for (int a = 0; a < 3; ++a) {
        float p;
        std::memcpy(&p, &node[a], 4);
        int e = int(int8_t(uint8_t(node[3] >> (8 * a))));
        float s = qscale(e);
        if (W == 4) {
            // the scale words the W = 4 kernels read; they must equal 2^e
            const uint32_t bits = a == 0 ? node[14] << 16 : (a == 1 ? node[14] & 0xFFFF0000u : node[15]);
            float sw;
            std::memcpy(&sw, &bits, 4);
            if (sw != s) s = std::numeric_limits<float>::quiet_NaN();  // fails every containment check
        }
        lo[a] = qdecode(q[(2 * a) * W + child], s, p);
        hi[a] = qdecode(q[(2 * a + 1) * W + child], s, p);
    }
}

}  // namespace kr
