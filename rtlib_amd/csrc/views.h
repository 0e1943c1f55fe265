// Plain-pointer views of device-resident data, passed by value to kernels.
#pragma once
#include <stdint.h>
#include "common.h"
#include "bvh.h"

#if !defined(__HIPCC__)
struct float4 { float x, y, z, w; };
struct uint2 { uint32_t x, y; };
struct uint4 { uint32_t x, y, z, w; };
#endif

namespace rtpg {

// Per-triangle shading word: material type in bits 0..3, light index + 1 in
// bits 4..31 (0 = not emissive).
constexpr uint32_t kTypeMask = 0xFu;
RTPG_HD uint32_t tri_type(uint32_t info) { return info & kTypeMask; }
RTPG_HD uint32_t tri_light(uint32_t info) { return (info >> 4) - 1u; }  // valid iff info >> 4 != 0
RTPG_HD bool tri_is_light(uint32_t info) { return (info >> 4) != 0u; }

struct SceneView {
  const BvhNode4* nodes;
  const TriRec* tris;
  const float4* shade;     // leaf order: albedo rgb, w = roughness / shininess / ior
  const float4* spec;      // leaf order: specular rgb (Phong), may be null
  const uint32_t* info;    // leaf order: type | (light + 1) << 4
  const float4* lights;    // 4 x float4 per emissive triangle: (v0, area) (e1, pdf_sel) (e2, tri) (Le, -)
  const float* light_cdf;  // inclusive CDF over lights
  uint32_t n_lights;
  uint32_t n_tris;
  uint32_t n_nodes;
  uint32_t max_stack;
};

// Flags of a render launch.  The low seven bits keep the meaning of the
// reference's ParamFlag (SimpleKernel.h:149-159).
enum RenderFlag : uint32_t {
  kFlagNee = 1u,
  kFlagRis = 2u,
  kFlagUseGrid = 4u,
  kFlagUseTree = 8u,
  kFlagGuideSample = 16u,  // sample directions from the guide (reference: BUILD)
  kFlagFinal = 32u,        // no recording (reference: FINAL)
  kFlagLocate = 64u,       // insert hash cells (reference: LOCATE)
  kFlagMis = 128u,
  kFlagRR = 256u,
  kFlagUseMlp = 512u,
};

struct Camera {
  float eye[3], u[3], v[3], w[3];
};

// SD-tree, device side (flattened; see sdtree.hip).
struct SdTreeView {
  const uint32_t* stree;   // bit 31 set: leaf, low bits DTree index; else first child
  const uint32_t* jump;    // 2^(3 jump_bits) node words after 3 jump_bits levels (or null)
  uint32_t jump_bits;
  float aabb_min[3];
  float aabb_size;         // cube edge
  const float4* info;      // per DTree: (samp_base bits, samp_sum, samp_sw, build_base bits)
  float* build_sw;         // per DTree, recorded statistical weight
  float4* build_sums;      // node pools
  const uint2* build_child;
  const float4* samp_sums;
  const uint2* samp_child;
  // sampling nodes interleaved for the tracer: per node (sums), (children, 0, 0)
  const uint4* samp_node;
  // pdf jump table: per DTree and cell of the 2^kPdfJumpBits square grid of
  // canonical directions, (bits of the density product of the first
  // kPdfJumpBits levels, node to continue from or kJumpDone, the DTree's
  // node base, 0); a DTree of mean 0 has product 1 and kJumpDone everywhere
  const uint4* samp_jump;
  // record jump table of the building trees, per DTree and cell: (x, base of
  // the DTree's nodes); x with bit 31 set: the descent ends within
  // kPdfJumpBits levels, at node << 2 | slot (low bits); else the node it
  // continues from
  const uint2* build_jump;
  float fraction;
  uint32_t n_dtrees;
  // recording filters: bit 0 spatial box, bit 1 directional box; per DTree the
  // depth of its STree leaf (root = 0), for the voxel size of the spatial box
  uint32_t filter;
  const uint32_t* leaf_depth;
};

constexpr uint32_t kPdfJumpBits = 4;
constexpr uint32_t kPdfJumpCells = 1u << (2 * kPdfJumpBits);
constexpr uint32_t kJumpDone = 0xFFFFFFFFu;

enum RecordFilter : uint32_t {
  kFilterSpatialBox = 1u,      // reference SpatialFilterBox
  kFilterDirectionalBox = 2u,  // reference DirectionalFilterBox
};

// Hash grid + Morton quadtrees (reference OPX7UtilsGrid.h / OPX7UtilsMorton.h).
struct HashGridView {
  float origin[3];
  float size[3];
  uint32_t bounds[3];
  uint32_t slots;          // multiple of 32
  uint32_t key_mode;       // 0: reference pcg1d chain, 1: 30-bit Morton key
  const uint32_t* prv;     // checksums of the previous pass (read-only lookups)
  uint32_t* cur;           // checksums being inserted
};

struct MortonView {
  float* build;            // slots * per_cell
  const float* samp;
  uint32_t level;
  uint32_t per_cell;
  float fraction;
};

struct RenderParams {
  uint32_t width, height;
  uint32_t flags;
  uint32_t max_depth;
  uint32_t rr_depth;
  uint32_t num_candidates;
  uint32_t spp;            // samples in this launch
  uint32_t sample_offset;  // index of the first sample (for the RNG)
  uint64_t seed;
  float bg[3];
  float fraction;
};

}  // namespace rtpg
