// Host-side SAH BVH4 builder and the device layout it produces.
//
// Replaces the OptiX GAS/IAS build of the reference (OPX7Natives.cpp:195-244,
// inputs RTLibExtOPX7Test.h:525-656): triangles in, an acceleration structure
// out.  The layout is chosen for gfx950: one 80-byte node holds the four
// child boxes as binary16 offsets from the node's lower corner, rounded
// outwards, so a lane fetches a node with five 16-byte loads and
// v_fma_mix_f32 turns each plane into a ray distance in one instruction; leaf
// triangles are stored pre-transformed for Moller-Trumbore (v0, e1, e2) in
// traversal order, 48 bytes each.
#pragma once
#include <stdint.h>
#include <vector>

namespace rtpg {

// A child of a node, as the traversal stacks it: an inner node is its index
// (< 2^31); a leaf is kLeafEntry | (triangles - 1) << 27 | first triangle
// (first < 2^27, at most 16 triangles).  Empty slots hold 0 and a box no ray
// enters.
constexpr uint32_t kLeafEntry = 0x80000000u;

// Child boxes as binary16 bits of their offsets from org: x[0..3] the lower
// and x[4..7] the upper x planes of children 0..3 (y, z alike).  Lower
// planes are rounded down and upper planes up, each one binary16 step
// further, so a node's boxes contain the exact ones and the ray distances
// computed from them keep a margin beyond their own rounding.  An empty slot
// has +inf for both planes, which no ray enters.
struct alignas(16) BvhNode4 {
  uint16_t x[8], y[8], z[8];
  uint32_t entry[4];   // the children as traversal entries
  float org[3];        // lower corner: the minimum of the children's lower planes
  uint32_t pad;
};
static_assert(sizeof(BvhNode4) == 80, "BVH4 node must be 80 bytes");

struct TriRec {        // 48 bytes, leaf order
  float v0[4];         // w = original primitive id (bit pattern)
  float e1[4];
  float e2[4];
};

struct BvhBuildResult {
  std::vector<BvhNode4> nodes;
  std::vector<TriRec> tris;
  std::vector<uint32_t> order;  // leaf-order slot -> original primitive id
  float world_lo[3], world_hi[3];
  uint32_t max_depth = 0;       // BVH4 levels
  uint32_t max_stack = 0;       // worst-case traversal stack occupancy
  uint32_t max_leaf = 0;
};

// binary16 +inf: the planes of an empty child slot.
constexpr uint16_t kHalfInf = 0x7C00u;
// Leaves hold at most 3 triangles: a wave tests one leaf per lane per round,
// so larger leaves keep the lanes with small ones waiting (measured on the
// 1.05M-triangle room: closest-hit traversal 5.7% faster than with 8).
constexpr uint32_t kMaxLeafTris = 3;

// verts: 3*nv floats, idx: 3*nt uint32.  Uses up to `threads` host threads.
BvhBuildResult build_bvh4(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                          int threads);

}  // namespace rtpg
