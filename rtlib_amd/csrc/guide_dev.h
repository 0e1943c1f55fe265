// Device-side guide structures: SD-tree (Muller 2017) and hash grid + Morton
// direction quadtrees.  Each function names the reference routine whose
// semantics it keeps.
#pragma once
#include "common.h"
#include "views.h"

namespace rtpg {

constexpr uint32_t kInvalidCell = 0xFFFFFFFFu;

// ---------------------------------------------------------------- SD-tree
// STreeT::GetDTreeWrapper (OPX7UtilsPathGuiding.h:639-648, descent :561-567).
// The split axis of a node is its depth mod 3 (root x, SubDivide :1240).
// Unlike the reference the child index is clamped to {0,1}, so a point on the
// upper face of the box cannot index past the pair.
//
// The first 3G levels are resolved with one load from a table over the 2^G
// cubed grid of the box (entry: the word of the node reached after 3G
// levels, or the leaf met before).  The G child indices taken along an axis
// are the bits of clamp(floor(q 2^G), 0, 2^G - 1), and G doublings of q minus
// those bits leave exactly q 2^G - cell (power-of-two scaling and the
// subtraction of an integer in [q 2^G - 1, q 2^G] are exact in float), so the
// descent continues from the same node with the same coordinates.
__device__ __forceinline__ uint32_t stree_lookup(const SdTreeView& g, V3 p) {
  float qx = (p.x - g.aabb_min[0]) / g.aabb_size;
  float qy = (p.y - g.aabb_min[1]) / g.aabb_size;
  float qz = (p.z - g.aabb_min[2]) / g.aabb_size;
  uint32_t w;
  if (g.jump) {
    const uint32_t n = 1u << g.jump_bits;
    const float sn = (float)n;
    float gx = qx * sn, gy = qy * sn, gz = qz * sn;
    float cx = fminf(fmaxf(floorf(gx), 0.0f), sn - 1.0f);
    float cy = fminf(fmaxf(floorf(gy), 0.0f), sn - 1.0f);
    float cz = fminf(fmaxf(floorf(gz), 0.0f), sn - 1.0f);
    w = g.jump[((uint32_t)cz * n + (uint32_t)cy) * n + (uint32_t)cx];
    qx = gx - cx;
    qy = gy - cy;
    qz = gz - cz;
  } else {
    w = g.stree[0];
  }
  uint32_t axis = 0;
  while (!(w & kLeafBit)) {
    float& q = axis == 0 ? qx : (axis == 1 ? qy : qz);
    float pa = q * 2.0f;
    int idx = (int)pa;
    idx = idx < 0 ? 0 : (idx > 1 ? 1 : idx);
    q = pa - (float)idx;
    w = g.stree[w + (uint32_t)idx];
    axis = axis == 2 ? 0 : axis + 1;
  }
  return w & ~kLeafBit;
}

// Stochastic form of the spatial box filter (SpatialFilterBox: TraceVertexT
// ::Impl<Box>::Record :806-813 -> STreeT::Record :649-655 -> STreeNodeT::Record
// :573-611).  The reference gives every leaf overlapped by a box the record's
// statistical weight times w_j / volume (w_j: overlap volume), the box being
// the voxel of the leaf that holds the vertex, centred on the vertex.  Here
// one point is drawn uniformly in that box (u0..u2 in [0,1)) and the whole
// record goes to the leaf holding the point: leaf j receives it with
// probability w_j / volume, so each leaf's expected weight and flux are the
// reference's, for one extra lookup instead of a subtree walk.  A point
// outside the tree's cube returns kInvalidCell (the record is dropped; the
// reference's overlaps there are empty).  A leaf at depth D (split axis =
// depth mod 3) has been halved ceil(D/3) times along x, ceil((D-1)/3) along y
// and floor(D/3) along z.
__device__ __forceinline__ uint32_t stree_box_jitter(const SdTreeView& g, V3 p, uint32_t cell, float u0, float u1,
                                                     float u2) {
#pragma clang fp contract(off)
  uint32_t d = g.leaf_depth[cell];
  float ex = ldexpf(g.aabb_size, -(int)((d + 2u) / 3u));
  float ey = ldexpf(g.aabb_size, -(int)((d + 1u) / 3u));
  float ez = ldexpf(g.aabb_size, -(int)(d / 3u));
  V3 q{p.x + (u0 - 0.5f) * ex, p.y + (u1 - 0.5f) * ey, p.z + (u2 - 0.5f) * ez};
  float qx = (q.x - g.aabb_min[0]) / g.aabb_size;
  float qy = (q.y - g.aabb_min[1]) / g.aabb_size;
  float qz = (q.z - g.aabb_min[2]) / g.aabb_size;
  bool inside = qx >= 0.0f && qx < 1.0f && qy >= 0.0f && qy < 1.0f && qz >= 0.0f && qz < 1.0f;
  return inside ? stree_lookup(g, q) : kInvalidCell;
}

RTPG_HD uint32_t dnode_child(uint2 c, int idx) {
  uint32_t w = idx < 2 ? c.x : c.y;
  return (w >> ((idx & 1) * 16)) & 0xFFFFu;
}

RTPG_HD float f4c(float4 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }

// DTreeNodeT::GetChildIdx (OPX7UtilsPathGuiding.h:69-82): bit0 = x >= .5, bit1 = y >= .5.
RTPG_HD int dtree_child_idx(float& px, float& py) {
  int r = 0;
  if (px < 0.5f) {
    px *= 2.0f;
  } else {
    px = 2.0f * px - 1.0f;
    r |= 1;
  }
  if (py < 0.5f) {
    py *= 2.0f;
  } else {
    py = 2.0f * py - 1.0f;
    r |= 2;
  }
  return r;
}

// DTreeT::GetMean on the sampling tree (:444-448).
RTPG_HD float dtree_mean(float sum, float sw) {
  if (sw <= 0.0f) return 0.0f;
  float factor = kInv2Pi / (2.0f * sw);
  return factor * sum;
}

// Stochastic form of the directional box filter (DirectionalFilterBox:
// DTreeT::Impl<Box>::RecordIrradiance :429-443 -> DTreeNodeT::Record
// :126-169).  The reference spreads the flux over the leaf slots overlapped by
// a square of edge 0.5^depth centred on (u, v), depth being that of the slot
// holding (u, v) (DTreeNodeT::GetDepth :322-333), each slot taking its share
// of the square's area.  Here (u, v) moves to a uniform point of the square
// (r0, r1 in [0,1)), so each slot's expected flux is the reference's.  Returns
// false when the point leaves [0,1)^2: the flux is then dropped, the
// statistical weight still counts (as in the reference).
__device__ __forceinline__ bool dtree_box_jitter(const SdTreeView& g, uint32_t cell, float& u, float& v, float r0,
                                                 float r1) {
#pragma clang fp contract(off)
  uint32_t base = fbits(g.info[cell].w);
  float pu = u, pv = v;
  int depth = 1;
  uint32_t node = 0;
  int idx = dtree_child_idx(pu, pv);
  for (;;) {
    uint32_t ch = dnode_child(g.build_child[base + node], idx);
    if (ch == 0) break;
    node = ch;
    idx = dtree_child_idx(pu, pv);
    ++depth;
  }
  float e = ldexpf(1.0f, -depth);
  u = u + (r0 - 0.5f) * e;
  v = v + (r1 - 0.5f) * e;
  return u >= 0.0f && u < 1.0f && v >= 0.0f && v < 1.0f;
}

// DTreeNodeT::Pdf (:278-316) times 1/(4 pi) as in DTreeT::Pdf (:484-489):
// density on the sphere of the sampling tree of DTree `cell`.
__device__ __forceinline__ float dtree_pdf(const SdTreeView& g, uint32_t cell, float px, float py) {
#pragma clang fp contract(off)
  float result = 1.0f;
  uint32_t node = 0, base;
  if (g.samp_jump) {
    // the first kPdfJumpBits levels (and the mean-0 case) from the jump
    // table: their child indices are the top bits of (u, v), and
    // kPdfJumpBits doublings of u minus those bits leave exactly
    // u 2^kPdfJumpBits - cell (as in stree_lookup)
    const float sn = (float)(1u << kPdfJumpBits);
    float gx = px * sn, gy = py * sn;
    float cx = fminf(fmaxf(floorf(gx), 0.0f), sn - 1.0f);
    float cy = fminf(fmaxf(floorf(gy), 0.0f), sn - 1.0f);
    px = gx - cx;
    py = gy - cy;
    uint4 e = g.samp_jump[(size_t)cell * kPdfJumpCells + ((uint32_t)cy << kPdfJumpBits) + (uint32_t)cx];
    result = bitsf(e.x);
    if (e.y == kJumpDone) return result * kInv2Pi / 2.0f;
    node = e.y;
    base = e.z;
  } else {
    float4 inf = g.info[cell];
    if (dtree_mean(inf.y, inf.z) <= 0.0f) return 1.0f * kInv2Pi / 2.0f;
    base = fbits(inf.x);
  }
  int idx = dtree_child_idx(px, py);
  for (;;) {
    uint4 r0 = g.samp_node[2u * (base + node)], r1 = g.samp_node[2u * (base + node) + 1u];
    float4 s = make_float4(bitsf(r0.x), bitsf(r0.y), bitsf(r0.z), bitsf(r0.w));
    float total = s.x + s.y + s.z + s.w;
    float si = f4c(s, idx);
    if (si <= 0.0f) {
      result = 0.0f;
      break;
    }
    result *= 4.0f * si / total;
    uint32_t ch = dnode_child(make_uint2(r1.x, r1.y), idx);
    if (ch == 0) break;
    node = ch;
    idx = dtree_child_idx(px, py);
  }
  return result * kInv2Pi / 2.0f;
}

// DTreeNodeT::Sample (:228-277) with the clamp of DTreeT::Sample (:477-482):
// canonical direction drawn from the sampling tree of DTree `cell`.  pdf: the
// sphere density of the drawn point as DTreeNodeT::SampleAndPdf (:178-227)
// accumulates it along the same descent (4 s_i / sum per level, 1 / (4 pi)).
__device__ __forceinline__ void dtree_sample(const SdTreeView& g, uint32_t cell, Pcg32& rng,
                                             float& ox, float& oy, float& pdf) {
#pragma clang fp contract(off)
  float4 inf = g.info[cell];
  if (dtree_mean(inf.y, inf.z) <= 0.0f) {
    ox = rnd(rng);
    oy = rnd(rng);
    pdf = 1.0f * kInv2Pi / 2.0f;
    return;
  }
  float result = 1.0f;
  uint32_t base = fbits(inf.x);
  uint32_t node = 0;
  float rx = 0.0f, ry = 0.0f, size = 1.0f;
  for (;;) {
    uint4 r0 = g.samp_node[2u * (base + node)], r1 = g.samp_node[2u * (base + node) + 1u];
    float4 s = make_float4(bitsf(r0.x), bitsf(r0.y), bitsf(r0.z), bitsf(r0.w));
    int idx = 0;
    float partial = s.x + s.z;
    float total = s.x + s.y + s.z + s.w;
    float boundary = partial / total;
    float ox0 = 0.0f, oy0 = 0.0f;
    float sample = rnd(rng);
    if (sample < boundary) {
      sample = sample / boundary;
      boundary = s.x / partial;
    } else {
      partial = total - partial;
      ox0 = 0.5f;
      sample = (sample - boundary) / (1.0f - boundary);
      boundary = s.y / partial;
      idx |= 1;
    }
    if (sample >= boundary) {
      oy0 = 0.5f;
      idx |= 2;
    }
    float t0 = size * ox0;
    float t1 = size * oy0;
    rx = rx + t0;
    ry = ry + t1;
    size *= 0.5f;
    float si = f4c(s, idx);
    result = si > 0.0f ? result * (4.0f * si / total) : 0.0f;
    uint32_t ch = dnode_child(make_uint2(r1.x, r1.y), idx);
    if (ch == 0 || si <= 0.0f) {
      float a = rnd(rng);
      float b = rnd(rng);
      float t2 = size * a;
      float t3 = size * b;
      rx = rx + t2;
      ry = ry + t3;
      break;
    }
    node = ch;
  }
  ox = fminf(fmaxf(rx, 0.0f), 1.0f);
  oy = fminf(fmaxf(ry, 0.0f), 1.0f);
  pdf = result * kInv2Pi / 2.0f;
}

// ---------------------------------------------------------------- hash grid
// DoubleBufferedHashGrid3T (OPX7UtilsGrid.h:157-353): cell coordinates, bucket
// and checksum.  key_mode 1 replaces the pcg1d chain by a 30-bit Morton key of
// 10-bit cell coordinates: bucket = pcg1d(key) % (slots/32), checksum = key+1.
RTPG_HD void grid_cell(const HashGridView& h, V3 p, uint32_t& ix, uint32_t& iy, uint32_t& iz) {
  float fx = (p.x - h.origin[0]) / h.size[0];
  float fy = (p.y - h.origin[1]) / h.size[1];
  float fz = (p.z - h.origin[2]) / h.size[2];
  float gx = (float)h.bounds[0] * fx;
  float gy = (float)h.bounds[1] * fy;
  float gz = (float)h.bounds[2] * fz;
  // static_cast<unsigned int> of a negative float is undefined in C++; the
  // reference's CUDA conversion saturates it to 0, which is what is kept here.
  uint32_t cx = gx > 0.0f ? (gx < 4294967040.0f ? (uint32_t)gx : 0xFFFFFFFFu) : 0u;
  uint32_t cy = gy > 0.0f ? (gy < 4294967040.0f ? (uint32_t)gy : 0xFFFFFFFFu) : 0u;
  uint32_t cz = gz > 0.0f ? (gz < 4294967040.0f ? (uint32_t)gz : 0xFFFFFFFFu) : 0u;
  ix = cx < h.bounds[0] - 1u ? cx : h.bounds[0] - 1u;
  iy = cy < h.bounds[1] - 1u ? cy : h.bounds[1] - 1u;
  iz = cz < h.bounds[2] - 1u ? cz : h.bounds[2] - 1u;
}

RTPG_HD void grid_hash(const HashGridView& h, uint32_t ix, uint32_t iy, uint32_t iz,
                       uint32_t& bucket, uint32_t& checksum) {
  uint32_t nb = h.slots / 32u;
  if (h.key_mode == 1u) {
    uint32_t key = morton3_30(ix, iy, iz);
    bucket = pcg1d(key) % nb;
    checksum = key + 1u;
  } else {
    bucket = pcg1d(h.slots + iz + pcg1d(iy + pcg1d(ix))) % nb;
    checksum = 1u + h.bounds[0] * h.bounds[1] * iz + h.bounds[0] * iy + ix;
  }
}

// GetPrvCellIndex (:298-319): read-only probe of the previous pass's table.
// The slots are scanned in order up to the first empty one; the whole
// 128-byte block is read at once (eight 16-byte loads issued together), and
// the scan runs on bit masks: the result is the first slot holding the
// checksum if no empty slot comes before it.
RTPG_HD uint32_t grid_find_prv(const HashGridView& h, uint32_t bucket, uint32_t checksum) {
  const uint4* blk = reinterpret_cast<const uint4*>(h.prv + 32u * bucket);
  uint4 v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = blk[q];
  uint32_t empty = 0u, hit = 0u;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t c[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      empty |= (c[j] == 0u ? 1u : 0u) << (4 * q + j);
      hit |= (c[j] == checksum ? 1u : 0u) << (4 * q + j);
    }
  }
  if (hit == 0u) return kInvalidCell;
  const uint32_t first_hit = (uint32_t)__builtin_ctz(hit);
  const uint32_t first_empty = empty ? (uint32_t)__builtin_ctz(empty) : 32u;
  return first_hit < first_empty ? 32u * bucket + first_hit : kInvalidCell;
}

#if defined(__HIPCC__)
// GetCurCellIndex (:266-297): insert-or-find with atomicCAS(0 -> checksum).
__device__ __forceinline__ uint32_t grid_find_cur(const HashGridView& h, uint32_t bucket,
                                                  uint32_t checksum) {
  for (uint32_t i = 0; i < 32u; ++i) {
    uint32_t prev = atomicCAS(&h.cur[32u * bucket + i], 0u, checksum);
    if (prev == 0u || prev == checksum) return 32u * bucket + i;
  }
  return kInvalidCell;
}
#endif

// ---------------------------------------------------------------- Morton quadtree
// Leaf offset of level L in the pyramid: (4^L - 1) / 3.
RTPG_HD uint32_t mq_level_offset(uint32_t level) { return ((1u << (2u * level)) - 1u) / 3u; }

// MortonQuadTreeT::Pdf (OPX7UtilsMorton.h:401-414) times 1/(4 pi) (wrapper :515-523).
__device__ __forceinline__ float mq_pdf(const MortonView& m, uint32_t cell, float px, float py) {
  const float* w = m.samp + (size_t)m.per_cell * cell;
  float root = w[0];
  if (root <= 0.0f) return 1.0f * 0.25f * kInvPi;
  uint32_t nn = 1u << m.level;
  uint32_t x = (uint32_t)(px * (float)nn);
  uint32_t y = (uint32_t)(py * (float)nn);
  x = x < nn ? x : nn - 1u;
  y = y < nn ? y : nn - 1u;
  uint32_t code = morton2(x, y);
  float leaf = w[mq_level_offset(m.level) + code];
  return leaf / root * (float)(nn * nn) * 0.25f * kInvPi;
}

// MortonQuadTreeT::SampleAndPdf (:286-344); returns canonical coordinates and
// the density on the sphere.
__device__ __forceinline__ void mq_sample(const MortonView& m, uint32_t cell, Pcg32& rng,
                                          float& ox, float& oy, float& pdf) {
#pragma clang fp contract(off)
  const float* w = m.samp + (size_t)m.per_cell * cell;
  float total = w[0];
  if (total <= 0.0f) {
    ox = rnd(rng);
    oy = rnd(rng);
    pdf = 1.0f * 0.25f * kInvPi;
    return;
  }
  uint32_t size = 4, offset = 1, code = 0;
  for (uint32_t i = 0; i < m.level; ++i) {
    float w0 = w[offset + (code << 2) + 0], w1 = w[offset + (code << 2) + 1];
    float w2 = w[offset + (code << 2) + 2], w3 = w[offset + (code << 2) + 3];
    uint32_t local = 0;
    float partial = w0 + w2;
    float boundary = partial / total;
    float sample = rnd(rng);
    if (sample < boundary) {
      sample /= boundary;
      boundary = w0 / partial;
    } else {
      partial = w1 + w3;
      sample = (sample - boundary) / (1.0f - boundary);
      boundary = w1 / partial;
      local |= 1u;
    }
    if (sample >= boundary) local |= 2u;
    total = local == 0 ? w0 : (local == 1 ? w1 : (local == 2 ? w2 : w3));
    code = (code << 2) + local;
    offset += size;
    size *= 4u;
  }
  uint32_t xi = compact1by1_16(code), yi = compact1by1_16(code >> 1);
  float nn = (float)(1u << m.level);
  float a = rnd(rng);
  float b = rnd(rng);
  ox = ((float)xi + a) / nn;
  oy = ((float)yi + b) / nn;
  pdf = total / w[0] * (nn * nn) * 0.25f * kInvPi;
}

}  // namespace rtpg
