// kr_layout.h -- data layout shared by the host builder and the gfx950 kernels.
//
// Everything the kernels read lives in flat arrays in HBM:
//
//  Wide BVH node (W = 2, 4 or 8 children), 16-byte aligned, NODE_FLOATS(W)
//  floats each:  lo.x[W] hi.x[W] lo.y[W] hi.y[W] lo.z[W] hi.z[W] ref[W] pad
//  A child ref is an inner node, a leaf (LEAF bit | (count-1) << 27 |
//  first triangle) or EMPTY. The builder (WideBvh) refers to inner nodes by
//  index; in device memory the ref is the node's offset in 16-byte units
//  (index * node bytes / 16), so the kernels address it with a shift.
//
//  Triangles in leaf order, 3 float4 each (48 B):
//    {v0.xyz, prim index bits}, {v1.xyz, global triangle index bits},
//    {v2.xyz, material bucket bits: min(shader type, 2)}
//
//  Per input triangle (global index, gathered only on a hit): 9 floats of
//  vertex normals and the owning primitive.
#pragma once
#include <cstdint>

namespace kr {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLeafCountShift = 27;
constexpr uint32_t kLeafFirstMask = (1u << kLeafCountShift) - 1u;
constexpr uint32_t kMaxLeafTris = 16;
constexpr uint32_t kSphereBit = 0x80000000u;  // hit code: sphere index
constexpr uint32_t kMissCode = 0xFFFFFFFFu;

constexpr uint32_t node_floats(uint32_t w) { return (7 * w + 3) / 4 * 4; }

// Compressed wide node (the default; KR_SCENE_FULL_NODES selects the full
// FP32 boxes above): child boxes quantised to 8 bits per
// plane against the node's own box, as in Ylitie, Karras and Laine, "Efficient
// Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs" (HPG 2017):
//   word 0..2  origin p (float xyz; the node box's lower corner)
//   word 3     per-axis scale exponents e (int8 x 3): scale s_a = 2^e_a
//   word 4..   child refs (W x uint32, same encoding as above)
//   then       6W bytes: lo_x[W] hi_x[W] lo_y[W] hi_y[W] lo_z[W] hi_z[W]
//   W = 4 only: word 14 = upper halves of s_x (low 16 bits) and s_y (high 16
//              bits), word 15 = s_z (FP32), the scales the kernels read
// A child's box is lo = fma(q_lo, s, p), hi = fma(q_hi, s, p) in FP32; the
// encoder rounds q_lo down and q_hi up until the decoded box contains the
// exact one, so traversal only ever visits a superset of nodes. 64 B for
// W = 4 (112 B uncompressed), 96 B for W = 8 (224 B). The W = 4 and W = 8
// kernels fold the decode into the slab test (qnode_enter in kr_device.h).
constexpr uint32_t qnode_bytes(uint32_t w) { return (16 + 4 * w + 6 * w + 15) / 16 * 16; }

inline uint32_t make_leaf(uint32_t first, uint32_t count) {
    return kLeafBit | ((count - 1) << kLeafCountShift) | first;
}

// Device-side primitive / shader / light records.
struct PrimRec {
    uint32_t type;       // KR_PRIM_*
    uint32_t shader;
    uint32_t tri_begin;  // mesh: first global triangle
    uint32_t sphere;     // sphere: index into sphere arrays
};

struct ShaderRec {
    uint32_t type;   // KR_SHADER_*
    uint32_t dist;   // KR_DIST_*
    float alpha2;    // roughness^2
    float roughness; // Beckmann roughness of GlossyBSDF / GlassBSDF / RefractionBSDF, else 0
    float cr[4];     // Lambert color*R; Emission (1,1,1); the others color
    float emission[4];
    float ior;       // GlassBSDF / RefractionBSDF
    uint32_t fac_type;  // Mixed: KR_SCALA_*
    float fac_value;    // Mixed: Constant value or Fresnel ior
    uint32_t child0, child1;  // Mixed: shader a, shader b (b when u < fac)
    uint32_t pad[3];
};

// Light kinds in LightRec::type: KR_PRIM_MESH / KR_PRIM_SPHERE area lights,
// and point lights.
constexpr uint32_t kLightPoint = 2;

struct LightRec {
    uint32_t prim;
    uint32_t type;      // KR_PRIM_*
    uint32_t table;     // mesh: first entry in the light table (1024 entries)
    float pdf;          // mesh: 1 / total padded area
    float emission[4];  // E = color * strength
    float center[4];    // sphere: centre.xyz, r2; point light: position.xyz
};

constexpr uint32_t kLightTable = 1024;

}  // namespace kr
