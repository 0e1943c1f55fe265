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
  // Optional per-triangle attributes, leaf order (null when absent).
  // tattr: 2 x float4 per triangle, (u0, v0, u1, v1) and (u2, v2, texture
  // words): word z = diffuse | specular << 16, word w = emission on hit |
  // emission when light-sampled << 16, 16-bit indices, kTexNone = factor 1.
  const float4* tattr;
  const float4* vnorm;     // 3 x float4 per triangle: the 9 vertex-normal floats in order
  const uint4* tex;        // per texture: first texel, width, height, wrap (1) / clamp (0)
  const uint32_t* texels;  // RGBA8, R in the low byte, rows bottom-up
};

constexpr uint32_t kTexNone = 0xFFFFu;

// Texture coordinate at barycentrics (b0, b1, b2) of vertices 0, 1, 2, in
// the reference's order of terms (GetTexCrd, SimpleKernel.h:70-77, 251-257).
RTPG_HD void tri_texcoord(float4 a, float4 b, float b0, float b1, float b2, float& u, float& v) {
#pragma clang fp contract(off)
  u = b0 * a.x + b1 * a.z + b2 * b.x;
  v = b0 * a.y + b1 * a.w + b2 * b.y;
}

// Bilinear fetch of an RGBA8 texture at normalised (u, v): the filtering of
// a CUDA texture object with linear filter, normalised coordinates and
// normalised-float reads (CUDA C Programming Guide, "Texture Fetching"):
// x = u W - 1/2, i = floor(x), alpha = frac(x) held with 8 fractional bits
// (rounded to nearest), texel indices wrapped (wrap mode, after u -> frac(u))
// or clamped to the edge, texel channels c / 255.  Returns rgb.
RTPG_HD V3 tex_sample(uint4 t, const uint32_t* texels, float u, float v) {
#pragma clang fp contract(off)
  if (!isfinite(u)) u = 0.0f;
  if (!isfinite(v)) v = 0.0f;
  if (t.w) {
    u = u - floorf(u);
    v = v - floorf(v);
  } else {
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
  }
  const int w = (int)t.y, h = (int)t.z;
  float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = rintf((x - fx) * 256.0f) * (1.0f / 256.0f);
  float b = rintf((y - fy) * 256.0f) * (1.0f / 256.0f);
  int i0 = (int)fx, j0 = (int)fy, i1 = i0 + 1, j1 = j0 + 1;
  if (t.w) {
    i0 = (i0 % w + w) % w;
    i1 = (i1 % w + w) % w;
    j0 = (j0 % h + h) % h;
    j1 = (j1 % h + h) % h;
  } else {
    i0 = i0 < 0 ? 0 : (i0 >= w ? w - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 >= w ? w - 1 : i1);
    j0 = j0 < 0 ? 0 : (j0 >= h ? h - 1 : j0);
    j1 = j1 < 0 ? 0 : (j1 >= h ? h - 1 : j1);
  }
  const uint32_t* base = texels + t.x;
  uint32_t c00 = base[(size_t)j0 * w + i0], c10 = base[(size_t)j0 * w + i1];
  uint32_t c01 = base[(size_t)j1 * w + i0], c11 = base[(size_t)j1 * w + i1];
  float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
  float r[3];
  for (int k = 0; k < 3; ++k) {
    int sh = 8 * k;
    float t00 = (float)((c00 >> sh) & 0xFFu) / 255.0f, t10 = (float)((c10 >> sh) & 0xFFu) / 255.0f;
    float t01 = (float)((c01 >> sh) & 0xFFu) / 255.0f, t11 = (float)((c11 >> sh) & 0xFFu) / 255.0f;
    r[k] = w00 * t00 + w10 * t10 + w01 * t01 + w11 * t11;
  }
  return V3{r[0], r[1], r[2]};
}

// Factor of texture slot `id` (kTexNone: 1).
RTPG_HD V3 tex_factor(const SceneView& sc, uint32_t id, float u, float v) {
  if (id == kTexNone) return V3{1.0f, 1.0f, 1.0f};
  return tex_sample(sc.tex[id], sc.texels, u, v);
}

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
  kFlagMoments = 1024u,  // per-pixel sample moments (no reference counterpart)
};

struct Camera {
  float eye[3], u[3], v[3], w[3];
};

// SD-tree, device side (flattened; see sdtree.hip).  Spatial tree node words:
// a leaf holds kLeafBit | its DTree, an inner node the index of the first of
// its two children.
constexpr uint32_t kLeafBit = 0x80000000u;

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

constexpr uint32_t kPdfJumpBits = 5;
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
