// Host check of the BVH builder and the compressed node encoding (built and
// run by tests/test_bvh_host.py with g++; no GPU):
//   bvh_check <tris.bin> <width>
// tris.bin: uint32 n, then n x 9 floats (v0, v1, v2). Verifies that every
// triangle sits in exactly one leaf, that every child box contains its
// subtree, and that every decoded compressed child box contains the exact
// box; prints "ok nodes=<n> depth=<d> qbytes=<b>" or the first violation.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "kr_layout.h"

using namespace kr;

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t n = 0;
    if (std::fread(&n, 4, 1, f) != 1) return 2;
    std::vector<float> tv(size_t(n) * 9);
    if (std::fread(tv.data(), 4, tv.size(), f) != tv.size()) return 2;
    std::fclose(f);
    uint32_t W = uint32_t(std::atoi(argv[2]));
    WideBvh bvh;
    std::string err;
    if (!build_wide_bvh(tv.data(), n, W, 4, bvh, err)) {
        std::printf("build failed: %s\n", err.c_str());
        return 1;
    }
    const uint32_t NF = node_floats(W);
    std::vector<int> seen(n, 0);
    // walk: every triangle of a leaf inside all boxes on its path
    struct Item {
        uint32_t ref;
        float lo[3], hi[3];
    };
    std::vector<Item> stack;
    Item root{0, {-1e30f, -1e30f, -1e30f}, {1e30f, 1e30f, 1e30f}};
    stack.push_back(root);
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        if (it.ref & kLeafBit) {
            uint32_t first = it.ref & kLeafFirstMask, cnt = ((it.ref >> kLeafCountShift) & 0xF) + 1;
            for (uint32_t k = 0; k < cnt; ++k) {
                uint32_t t = bvh.tri_order[first + k];
                seen[t] += 1;
                for (int v = 0; v < 3; ++v)
                    for (int a = 0; a < 3; ++a) {
                        float x = tv[9 * size_t(t) + 3 * v + a];
                        if (x < it.lo[a] || x > it.hi[a]) {
                            std::printf("triangle %u outside its leaf box\n", t);
                            return 1;
                        }
                    }
            }
            continue;
        }
        const float *nd = bvh.nodes.data() + size_t(it.ref) * NF;
        const uint32_t *ref = reinterpret_cast<const uint32_t *>(nd + 6 * W);
        for (uint32_t c = 0; c < W; ++c) {
            if (ref[c] == kEmpty) continue;
            Item ch{ref[c], {}, {}};
            for (int a = 0; a < 3; ++a) {
                ch.lo[a] = nd[(2 * a) * W + c];
                ch.hi[a] = nd[(2 * a + 1) * W + c];
            }
            stack.push_back(ch);
        }
    }
    for (uint32_t t = 0; t < n; ++t)
        if (seen[t] != 1) {
            std::printf("triangle %u in %d leaves\n", t, seen[t]);
            return 1;
        }
    std::vector<uint32_t> q;
    compress_wide_nodes(bvh, q);
    const uint32_t NW = qnode_bytes(W) / 4;
    for (uint32_t i = 0; i < bvh.num_nodes; ++i) {
        const float *nd = bvh.nodes.data() + size_t(i) * NF;
        const uint32_t *ref = reinterpret_cast<const uint32_t *>(nd + 6 * W);
        for (uint32_t c = 0; c < W; ++c) {
            if (ref[c] == kEmpty) continue;
            if (q[size_t(i) * NW + 4 + c] != ref[c]) {
                std::printf("node %u child %u: ref changed\n", i, c);
                return 1;
            }
            float lo[3], hi[3];
            decode_qnode_box(q.data() + size_t(i) * NW, W, c, lo, hi);
            for (int a = 0; a < 3; ++a)
                if (!(lo[a] <= nd[(2 * a) * W + c] && hi[a] >= nd[(2 * a + 1) * W + c])) {
                    std::printf("node %u child %u axis %d: decoded box [%.9g, %.9g] misses [%.9g, %.9g]\n", i, c, a,
                                lo[a], hi[a], nd[(2 * a) * W + c], nd[(2 * a + 1) * W + c]);
                    return 1;
                }
        }
    }
    std::printf("ok nodes=%u depth=%u qbytes=%u\n", bvh.num_nodes, bvh.depth, qnode_bytes(W));
    return 0;
}
