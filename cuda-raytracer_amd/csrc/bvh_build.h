// bvh_build.h -- host builder of the wide BVH the gfx950 traversal kernels read.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace kr {

struct WideBvh {
    uint32_t width = 4;
    std::vector<float> nodes;          // node_floats(width) floats per node
    std::vector<uint32_t> tri_order;   // leaf order -> input triangle index
    uint32_t num_nodes = 0, num_leaves = 0, depth = 0;
};

// Binned-SAH binary build (32 bins on each axis, leaves of at most leaf_size triangles)
// collapsed to `width` children per node by repeatedly opening the child of
// largest surface area. tri_verts holds 9 floats (v0, v1, v2) per triangle.
bool build_wide_bvh(const float *tri_verts, uint32_t num_tris, uint32_t width, uint32_t leaf_size, WideBvh &out,
                    std::string &err);

// Re-encodes the nodes of `bvh` in the compressed layout (qnode_bytes(width)
// bytes each, conservative 8-bit child boxes).
void compress_wide_nodes(const WideBvh &bvh, std::vector<uint32_t> &out);

// Decoded child box of a compressed node, exactly as the kernels decode it.
void decode_qnode_box(const uint32_t *node, uint32_t width, uint32_t child, float lo[3], float hi[3]);

}  // namespace kr
