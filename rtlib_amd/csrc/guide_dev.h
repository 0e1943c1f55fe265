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

// One density walk of a DTree's sampling tree (DTreeNodeT::Pdf :278-316,
// times 1/(4 pi) as in DTreeT::Pdf :484-489): the state kept between the
// loads of its descent, so that two descents can share their load rounds.
struct DPdfWalk {
  uint32_t base, node;
  float result, px, py;
  int idx;
  bool live;
};

// Start of a walk from the first word of its DTree: with the jump table, the
// entry j read at (px, py) (first kPdfJumpBits levels and the mean-0 case:
// their child indices are the top bits of (u, v), and kPdfJumpBits doublings
// of u minus those bits leave exactly u 2^kPdfJumpBits - cell, as in
// stree_lookup; dpdf_jump_slot has already rescaled px, py); without it, the
// DTree's info word.
RTPG_HD DPdfWalk dpdf_begin(bool jump, uint4 j, float px, float py) {
  DPdfWalk w;
  w.node = 0u;
  w.px = px;
  w.py = py;
  w.idx = 0;
  if (jump) {
    w.result = bitsf(j.x);
    w.live = j.y != kJumpDone;
    w.node = w.live ? j.y : 0u;
    w.base = j.z;
  } else {
    w.result = 1.0f;
    w.live = !(dtree_mean(bitsf(j.y), bitsf(j.z)) <= 0.0f);
    w.base = j.x;
  }
  if (w.live) w.idx = dtree_child_idx(w.px, w.py);
  return w;
}

// Index of the jump entry of (px, py) in DTree `cell`; rescales px, py.
RTPG_HD size_t dpdf_jump_slot(uint32_t cell, float& px, float& py) {
#pragma clang fp contract(off)
  const float sn = (float)(1u << kPdfJumpBits);
  float gx = px * sn, gy = py * sn;
  float cx = fminf(fmaxf(floorf(gx), 0.0f), sn - 1.0f);
  float cy = fminf(fmaxf(floorf(gy), 0.0f), sn - 1.0f);
  px = gx - cx;
  py = gy - cy;
  return (size_t)cell * kPdfJumpCells + ((uint32_t)cy << kPdfJumpBits) + (uint32_t)cx;
}

// One level of a walk, given its node's sums r0 and child words r1.
RTPG_HD void dpdf_step(DPdfWalk& w, uint4 r0, uint2 r1) {
#pragma clang fp contract(off)
  float4 s = make_float4(bitsf(r0.x), bitsf(r0.y), bitsf(r0.z), bitsf(r0.w));
  float total = s.x + s.y + s.z + s.w;
  float si = f4c(s, w.idx);
  if (si <= 0.0f) {
    w.result = 0.0f;
    w.live = false;
    return;
  }
  w.result *= 4.0f * si / total;
  uint32_t ch = dnode_child(r1, w.idx);
  if (ch == 0) {
    w.live = false;
    return;
  }
  w.node = ch;
  w.idx = dtree_child_idx(w.px, w.py);
}

// Two descents of DTree `cell`'s sampling tree in one loop: every round
// issues the node loads of both before either is used, so a shading point
// waits for the longer descent rather than for the sum of the two.
//  - chain a, with a_sample set: DTreeNodeT::Sample (:228-277) with the clamp
//    of DTreeT::Sample (:477-482): (ax, ay) <- a canonical direction, apdf <-
//    the sphere density of the drawn point as DTreeNodeT::SampleAndPdf
//    (:178-227) accumulates it along the same descent (4 s_i / sum per level,
//    1 / (4 pi)); one random number per level and two at the end, in order;
//  - chain a, without a_sample: apdf <- the density at (ax, ay);
//  - chain b, when b_on: bpdf <- the density at (bx, by).
// Each chain performs the operations of the descent it stands for in the
// same order, so its result does not depend on the other chain.
__device__ __forceinline__ void dtree_pair(const SdTreeView& g, uint32_t cell, bool a_sample, Pcg32& rng,
                                           float& ax, float& ay, float& apdf, bool b_on, float bx, float by,
                                           float& bpdf) {
#pragma clang fp contract(off)
  const bool jump = g.samp_jump != nullptr;
  const uint4* info = reinterpret_cast<const uint4*>(g.info + cell);
  float apx = ax, apy = ay;
  const uint4* pa = (a_sample || !jump) ? info : g.samp_jump + dpdf_jump_slot(cell, apx, apy);
  const uint4* pb = jump ? g.samp_jump + dpdf_jump_slot(cell, bx, by) : info;
  uint4 ja = *pa;
  uint4 jb = b_on ? *pb : make_uint4(0u, kJumpDone, 0u, 0u);
  DPdfWalk b = dpdf_begin(jump, jb, bx, by);
  b.live = b.live && b_on;
  // chain a: a density walk, or the sampling descent's state
  DPdfWalk a = dpdf_begin(a_sample ? false : jump, ja, apx, apy);
  float rx = 0.0f, ry = 0.0f, size = 1.0f;
  bool uniform = false;
  if (a_sample) {
    a.idx = 0;
    a.result = 1.0f;
    if (!a.live) {
      // an empty DTree samples the square uniformly
      ax = rnd(rng);
      ay = rnd(rng);
      uniform = true;
    }
  }
  while (a.live || b.live) {
    uint4 a0 = make_uint4(0u, 0u, 0u, 0u), b0 = a0;
    uint2 a1 = make_uint2(0u, 0u), b1 = a1;
    if (a.live) {
      a0 = g.samp_node[2u * (a.base + a.node)];
      a1 = *reinterpret_cast<const uint2*>(g.samp_node + 2u * (a.base + a.node) + 1u);
    }
    if (b.live) {
      b0 = g.samp_node[2u * (b.base + b.node)];
      b1 = *reinterpret_cast<const uint2*>(g.samp_node + 2u * (b.base + b.node) + 1u);
    }
    if (b.live) dpdf_step(b, b0, b1);
    if (a.live) {
      if (!a_sample) {
        dpdf_step(a, a0, a1);
      } else {
        float4 s = make_float4(bitsf(a0.x), bitsf(a0.y), bitsf(a0.z), bitsf(a0.w));
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
        a.result = si > 0.0f ? a.result * (4.0f * si / total) : 0.0f;
        uint32_t ch = dnode_child(a1, idx);
        if (ch == 0 || si <= 0.0f) {
          float u = rnd(rng);
          float v = rnd(rng);
          float t2 = size * u;
          float t3 = size * v;
          rx = rx + t2;
          ry = ry + t3;
          a.live = false;
        } else {
          a.node = ch;
        }
      }
    }
  }
  if (a_sample && !uniform) {
    ax = fminf(fmaxf(rx, 0.0f), 1.0f);
    ay = fminf(fmaxf(ry, 0.0f), 1.0f);
  }
  apdf = a.result * kInv2Pi / 2.0f;
  bpdf = b.result * kInv2Pi / 2.0f;
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
// The slots are scanned in order up to the first empty one, on bit masks:
// the result is the first slot holding the checksum if no empty slot comes
// before it.  The first 8 slots of the 128-byte block are read first (two
// 16-byte loads); the other 24 only when those hold neither the checksum nor
// an empty slot (a bucket holds about one cell on average).
RTPG_HD uint32_t grid_find_prv(const HashGridView& h, uint32_t bucket, uint32_t checksum) {
  const uint4* blk = reinterpret_cast<const uint4*>(h.prv + 32u * bucket);
  uint32_t empty = 0u, hit = 0u;
  auto scan = [&](int q0, int q1) {
    uint4 v[6];
    for (int q = q0; q < q1; ++q) v[q - q0] = blk[q];
    for (int q = q0; q < q1; ++q) {
      const uint32_t c[4] = {v[q - q0].x, v[q - q0].y, v[q - q0].z, v[q - q0].w};
      for (int j = 0; j < 4; ++j) {
        empty |= (c[j] == 0u ? 1u : 0u) << (4 * q + j);
        hit |= (c[j] == checksum ? 1u : 0u) << (4 * q + j);
      }
    }
  };
  scan(0, 2);
  if ((empty | hit) == 0u) scan(2, 8);
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

// A point of [0,1) from an index (the top 24 bits of i * mult), for the
// self-checks of the sampling entry points.
RTPG_HD float check_coord(uint32_t i, uint32_t mult) { return (float)((i * mult) >> 8) * 0x1p-24f; }

// Four floats at a 4-byte-aligned address (one 16-byte load).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

// Leaf bin of canonical (px, py) in a level-`level` Morton pyramid.
RTPG_HD uint32_t mq_leaf(uint32_t level, float px, float py) {
  uint32_t nn = 1u << level;
  uint32_t x = (uint32_t)(px * (float)nn);
  uint32_t y = (uint32_t)(py * (float)nn);
  x = x < nn ? x : nn - 1u;
  y = y < nn ? y : nn - 1u;
  return mq_level_offset(level) + morton2(x, y);
}

// Two uses of cell `cell`'s pyramid with their loads issued together (the
// root, both leaves and the first sampling level in one round):
//  - chain a, with a_sample set: MortonQuadTreeT::SampleAndPdf (:286-344),
//    (ax, ay) <- canonical coordinates, apdf <- the density on the sphere of
//    the sampled bin; one random number per level and two at the end;
//  - chain a, without a_sample: MortonQuadTreeT::Pdf (:401-414) at (ax, ay)
//    times 1/(4 pi) (wrapper :515-523);
//  - chain b, when b_on: the same density at (bx, by).
__device__ __forceinline__ void mq_pair(const MortonView& m, uint32_t cell, bool a_sample, Pcg32& rng, float& ax,
                                        float& ay, float& apdf, bool b_on, float bx, float by, float& bpdf) {
#pragma clang fp contract(off)
  const float* w = m.samp + (size_t)m.per_cell * cell;
  const uint32_t nn = 1u << m.level;
  const float root = w[0];
  const float lb = b_on ? w[mq_leaf(m.level, bx, by)] : 0.0f;
  const float la = a_sample ? 0.0f : w[mq_leaf(m.level, ax, ay)];
  // the four values of a pyramid node are consecutive floats: one 16-byte
  // load (4-byte aligned, as a cell's pyramid is 85 floats)
  f4u s1 = {0.0f, 0.0f, 0.0f, 0.0f};
  if (a_sample && m.level > 0u) s1 = *reinterpret_cast<const f4u*>(w + 1);
  bpdf = root <= 0.0f ? 1.0f * 0.25f * kInvPi : lb / root * (float)(nn * nn) * 0.25f * kInvPi;
  if (!a_sample) {
    apdf = root <= 0.0f ? 1.0f * 0.25f * kInvPi : la / root * (float)(nn * nn) * 0.25f * kInvPi;
    return;
  }
  float total = root;
  if (total <= 0.0f) {
    ax = rnd(rng);
    ay = rnd(rng);
    apdf = 1.0f * 0.25f * kInvPi;
    return;
  }
  uint32_t size = 4, offset = 1, code = 0;
  for (uint32_t i = 0; i < m.level; ++i) {
    const f4u q = i > 0 ? *reinterpret_cast<const f4u*>(w + offset + (code << 2)) : s1;
    const float w0 = q.x, w1 = q.y, w2 = q.z, w3 = q.w;
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
  float fn = (float)nn;
  float u = rnd(rng);
  float v = rnd(rng);
  ax = ((float)xi + u) / fn;
  ay = ((float)yi + v) / fn;
  apdf = total / root * (fn * fn) * 0.25f * kInvPi;
}

}  // namespace rtpg
