// Host-side SAH BVH4 builder and the device layout it produces.
//
// Replaces the OptiX GAS/IAS build of the reference (OPX7Natives.cpp:195-244,
// inputs RTLibExtOPX7Test.h:525-656): triangles in, an acceleration structure
// out.  The layout is chosen for gfx950: one 128-byte node (two 64-byte
// lines) holds the four child boxes in SoA order so a wave fetches a node
// with eight 16-byte loads, and leaf triangles are stored pre-transformed for
// Moller-Trumbore (v0, e1, e2) in traversal order, 48 bytes each.
#pragma once
#include <stdint.h>
#include <vector>

namespace rtpg {

// A child of a node, as the traversal stacks it: an inner node is its index
// (< 2^31); a leaf is kLeafEntry | (triangles - 1) << 27 | first triangle
// (first < 2^27, at most 16 triangles).  Empty slots hold 0 and a box no ray
// enters.
constexpr uint32_t kLeafEntry = 0x80000000u;

struct alignas(64) BvhNode4 {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  uint32_t entry[4];   // the children as traversal entries
  uint32_t pad[4];     // keeps the node at two 64-byte lines
};
static_assert(sizeof(BvhNode4) == 128, "BVH4 node must be 128 bytes");

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

// Empty child slots hold a degenerate box far away (1e30) that no ray
// within t < 1e20 can enter.
constexpr float kEmptyBox = 1.0e30f;
// Leaves hold at most 3 triangles: a wave tests one leaf per lane per round,
// so larger leaves keep the lanes with small ones waiting (measured on the
// 1.05M-triangle room: closest-hit traversal 5.7% faster than with 8).
constexpr uint32_t kMaxLeafTris = 3;

// verts: 3*nv floats, idx: 3*nt uint32.  Uses up to `threads` host threads.
BvhBuildResult build_bvh4(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                          int threads);

}  // namespace rtpg
