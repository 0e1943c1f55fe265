// Device kernels of the SD-tree: Build (post-order sums + area), Reset of the
// building trees from the sampling trees, sampling-tree copies after a
// spatial split, and record/sample/pdf/lookup entry kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "guide_dev.h"
#include "sdtree.h"

namespace rtpg {

static inline uint32_t nblk(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

// RTDTreeT::Build (OPX7UtilsPathGuiding.h:948-957) -> DTreeNodeT::Build
// (:356-369) and GetArea (:342-355).  Children always carry larger indices
// than their parent (DFS allocation in Reset), so one reverse sweep over a
// DTree's nodes visits every child before its parent and reproduces the
// recursive sums with the same addition order.
__global__ void k_dtree_build(DTreeArrays a, float* area) {
#pragma clang fp contract(off)
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.n) return;
  uint32_t base = a.b_base[j], cnt = a.b_cnt[j];
  for (uint32_t k = cnt; k-- > 0;) {
    uint2 c = a.b_child[base + k];
    float4 s = a.b_sums[base + k];
    float sv[4] = {s.x, s.y, s.z, s.w};
    for (int i = 0; i < 4; ++i) {
      uint32_t ch = dnode_child(c, i);
      if (ch != 0) {
        float4 cs = a.b_sums[base + ch];
        float sum = 0.0f;
        sum += cs.x;
        sum += cs.y;
        sum += cs.z;
        sum += cs.w;
        sv[i] = sum;
      }
    }
    float ar = 0.0f;
    for (int i = 0; i < 4; ++i) {
      if (sv[i] > 0.0f) {
        uint32_t ch = dnode_child(c, i);
        if (ch == 0) {
          ar += 1.0f / 4.0f;
        } else {
          ar += area[base + ch] / 4.0f;
        }
      }
    }
    a.b_sums[base + k] = make_float4(sv[0], sv[1], sv[2], sv[3]);
    area[base + k] = ar;
  }
  float4 r = a.b_sums[base];
  float sum = 0.0f;
  sum += r.x;
  sum += r.y;
  sum += r.z;
  sum += r.w;
  a.b_sum[j] = sum;
  a.b_area[j] = area[base];
}

__global__ void k_dtree_publish(DTreeArrays a) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.n) return;
  a.info[j] = make_float4(bitsf(a.s_base[j]), a.s_sum[j], a.s_sw[j], bitsf(a.b_base[j]));
}

// One 32-byte record per sampling node, so a descent step touches one line.
__global__ void k_dtree_interleave(DTreeArrays a) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.s_nodes) return;
  float4 s = a.s_sums[i];
  uint2 c = a.s_child[i];
  a.s_node[2 * i] = make_uint4(fbits(s.x), fbits(s.y), fbits(s.z), fbits(s.w));
  a.s_node[2 * i + 1] = make_uint4(c.x, c.y, 0u, 0u);
}

// Pdf jump table of the sampling trees: work-item (DTree j, cell c) runs the
// first kPdfJumpBits levels of the density walk (dpdf_step) for the directions of cell
// c (child index bits = the cell's coordinate bits, most significant first),
// multiplying the density factors in the same order, and records where the
// descent continues (kJumpDone when it ended within those levels).
__global__ void k_dtree_pdf_jump(DTreeArrays a) {
#pragma clang fp contract(off)
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n * kPdfJumpCells) return;
  uint32_t j = t / kPdfJumpCells, c = t % kPdfJumpCells;
  uint32_t cx = c % (1u << kPdfJumpBits), cy = c / (1u << kPdfJumpBits);
  uint32_t base = a.s_base[j];
  float result = 1.0f;
  uint32_t node = 0, cont = kJumpDone;
  float4 inf = a.info[j];
  // a DTree of mean 0 samples uniformly: density 1 / (4 pi) (DTreeT::Pdf)
  const uint32_t levels = dtree_mean(inf.y, inf.z) <= 0.0f ? 0u : kPdfJumpBits;
  for (uint32_t l = 0; l < levels; ++l) {
    uint32_t b = kPdfJumpBits - 1u - l;
    int idx = (int)(((cx >> b) & 1u) | (((cy >> b) & 1u) << 1));
    uint4 r0 = a.s_node[2u * (base + node)], r1 = a.s_node[2u * (base + node) + 1u];
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
    if (l + 1u == kPdfJumpBits) cont = node;
  }
  a.s_jump[t] = make_uint4(fbits(result), cont, base, 0u);
}

// Record jump table of the building trees: work-item (DTree j, cell c) runs
// the first kPdfJumpBits levels of the record descent (DTreeNodeT::Record
// :113-125) for the directions of cell c.
__global__ void k_dtree_record_jump(DTreeArrays a) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n * kPdfJumpCells) return;
  uint32_t j = t / kPdfJumpCells, c = t % kPdfJumpCells;
  uint32_t cx = c % (1u << kPdfJumpBits), cy = c / (1u << kPdfJumpBits);
  uint32_t base = a.b_base[j];
  uint32_t node = 0, e = 0;
  for (uint32_t l = 0; l < kPdfJumpBits; ++l) {
    uint32_t b = kPdfJumpBits - 1u - l;
    int idx = (int)(((cx >> b) & 1u) | (((cy >> b) & 1u) << 1));
    uint32_t ch = dnode_child(a.b_child[base + node], idx);
    if (ch == 0) {
      e = 0x80000000u | (node << 2) | (uint32_t)idx;
      break;
    }
    node = ch;
    e = node;
  }
  a.b_jump[t] = make_uint2(e, base);
}

// RTDTreeT::Reset (OPX7UtilsPathGuiding.h:888-947), one work-item per DTree.
// The DFS keeps the reference's LIFO order, so nodes get the reference's
// numbers.  A stack entry either copies a node of the source sampling tree
// (src != kSelf) or continues below a source leaf (src == kSelf), whose four
// sums are all equal to `val` (a quarter of the parent slot).
constexpr uint32_t kSelf = 0xFFFFFFFFu;
constexpr int kResetStack = 64;

__global__ void k_dtree_reset(DTreeArrays src, DTreeArrays dst, const uint32_t* map, uint32_t max_depth,
                              float th, int pass, uint32_t* err) {
#pragma clang fp contract(off)
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= dst.n) return;
  uint32_t sj = map ? map[j] : j;
  uint32_t sbase = src.s_base[sj];
  float total = src.s_sum[sj];
  uint32_t obase = pass ? dst.b_base[j] : 0u;
  uint32_t st_dst[kResetStack], st_src[kResetStack], st_depth[kResetStack];
  float st_val[kResetStack];
  int sp = 0;
  st_dst[0] = 0;
  st_src[0] = 0;
  st_depth[0] = 1;
  st_val[0] = 0.0f;
  sp = 1;
  uint32_t count = 1;
  uint32_t maxd = 0;
  if (pass) {
    dst.b_child[obase] = make_uint2(0u, 0u);
    dst.b_sums[obase] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  bool stop = false;
  while (sp > 0 && !stop) {
    --sp;
    uint32_t e_dst = st_dst[sp], e_src = st_src[sp], e_depth = st_depth[sp];
    float e_val = st_val[sp];
    maxd = e_depth > maxd ? e_depth : maxd;
    float4 ss = make_float4(e_val, e_val, e_val, e_val);
    uint2 sc = make_uint2(0u, 0u);
    if (e_src != kSelf) {
      ss = src.s_sums[sbase + e_src];
      sc = src.s_child[sbase + e_src];
    }
    uint32_t newc[4] = {0u, 0u, 0u, 0u};
    for (int i = 0; i < 4; ++i) {
      float si = f4c(ss, i);
      float fraction;
      if (total > 0.0f) {
        fraction = si / total;
      } else {
        // std::pow(0.25f, depth) evaluates in double; 0.25^d is exact in float
        // for every depth that can occur.
        fraction = ldexpf(1.0f, -2 * (int)e_depth);
      }
      if (e_depth < max_depth && fraction > th) {
        if (sp >= kResetStack) {
          atomicOr(err, 2u);
          stop = true;
          break;
        }
        uint32_t sch = (e_src != kSelf) ? dnode_child(sc, i) : 0u;
        if (e_src != kSelf && sch != 0u) {
          st_src[sp] = sch;
          st_val[sp] = 0.0f;
        } else {
          st_src[sp] = kSelf;
          st_val[sp] = si / 4.0f;
        }
        st_dst[sp] = count;
        st_depth[sp] = e_depth + 1u;
        ++sp;
        newc[i] = count;
        if (pass) {
          dst.b_child[obase + count] = make_uint2(0u, 0u);
          dst.b_sums[obase + count] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        ++count;
        if (count > 65535u) {
          stop = true;
          break;
        }
      }
    }
    if (pass) {
      dst.b_child[obase + e_dst] = make_uint2(newc[0] | (newc[1] << 16), newc[2] | (newc[3] << 16));
    }
  }
  if (!pass) {
    dst.b_cnt[j] = count;
  } else {
    dst.b_depth[j] = maxd;
    dst.b_sw[j] = 0.0f;
    dst.b_sum[j] = 0.0f;
    dst.b_area[j] = 0.0f;
  }
}

// Sampling tree of new DTree j := sampling tree of map[j] (SubDivide copies the
// whole wrapper, RTSTreeT::SubDivide :1231-1247).  One block per DTree.
__global__ void k_dtree_copy_sampling(DTreeArrays src, DTreeArrays dst, const uint32_t* map) {
  uint32_t j = blockIdx.x;
  if (j >= dst.n) return;
  uint32_t sj = map[j];
  uint32_t sb = src.s_base[sj], cnt = src.s_cnt[sj], db = dst.s_base[j];
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
    dst.s_sums[db + k] = src.s_sums[sb + k];
    dst.s_child[db + k] = src.s_child[sb + k];
  }
  if (threadIdx.x == 0) {
    dst.s_sw[j] = src.s_sw[sj];
    dst.s_sum[j] = src.s_sum[sj];
    dst.s_area[j] = src.s_area[sj];
    dst.s_depth[j] = src.s_depth[sj];
  }
}

// Three-phase exclusive scan (1024-element tiles).
constexpr uint32_t kScanTile = 1024;

__global__ void k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tile_sums) {
  __shared__ uint32_t s[kScanTile];
  uint32_t t = threadIdx.x;
  uint32_t g = blockIdx.x * kScanTile + t;
  uint32_t v = g < n ? in[g] : 0u;
  s[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < kScanTile; off <<= 1) {
    uint32_t x = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  if (g < n) out[g] = s[t] - v;
  if (t == kScanTile - 1) tile_sums[blockIdx.x] = s[t];
}

__global__ void k_scan_sums(uint32_t* tile_sums, uint32_t ntiles, uint32_t* total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t acc = 0;
  for (uint32_t i = 0; i < ntiles; ++i) {
    uint32_t v = tile_sums[i];
    tile_sums[i] = acc;
    acc += v;
  }
  *total = acc;
}

__global__ void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* tile_sums) {
  uint32_t g = blockIdx.x * kScanTile + threadIdx.x;
  if (g < n) out[g] += tile_sums[blockIdx.x];
}

// ------------------------------------------------------------------ STree refine
// RTSTreeT::Refine (OPX7UtilsPathGuiding.h:1275-1311) and RTSTree2T's
// SubDivide (:2062-2077) on the device.  The reference walks the tree depth
// first from a stack (the second child is visited first) and appends the
// children of every leaf it splits at the end of the node array, so:
//  - a leaf of building weight w > th splits, and its children inherit w / 2
//    (w / 8 for the octree) and are visited next: the leaf becomes a complete
//    subtree of k splits, b^k leaves (b = 2 or 8) and 2 (b^k - 1) new nodes
//    (a binary split adds 2 nodes and 1 leaf, an octree split 14 and 7);
//  - the new nodes of the leaves form one block per split leaf, the blocks in
//    the walk's order of the old leaves.
// With the old leaves listed in that order (order[q]), an exclusive prefix
// sum L of b^k gives every block's start, N_old + 2 (L[q] - q), and the
// position of every new leaf in the next walk order; a second prefix sum over
// the leaf flags of the new node array numbers the DTrees in node order
// (RTSTreeWrapperT::Upload :1480-1512).

// Bytes of RTSTreeWrapperT::GetApproxMemoryFootPrint summed over the nodes: a
// leaf's wrapper holds its two DTrees (24 bytes a node and 48), an inner node
// a default wrapper of two one-node trees (96 bytes).
__global__ void k_stree_footprint(const uint32_t* stree, uint32_t n, const uint32_t* b_cnt, const uint32_t* s_cnt,
                                  unsigned long long* stat) {
  __shared__ unsigned long long part[256];
  unsigned long long acc = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t w = stree[i];
    if (w & kLeafBit) {
      uint32_t d = w & ~kLeafBit;
      acc += 24ull * ((unsigned long long)b_cnt[d] + s_cnt[d]) + 48ull;
    } else {
      acc += 96ull;
    }
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(stat + kStatFootprint, part[0]);
}

// Split count k of the q-th leaf in walk order, b^k, and the sums the host
// sizes the new arrays from: leaves, building nodes (b^k copies of the leaf's
// reset tree) and sampling nodes (b^k copies of its sampling tree).
__global__ void k_stree_split_count(const uint32_t* stree, const uint32_t* order, uint32_t n_leaves,
                                    const float* b_sw, const uint32_t* reset_cnt, const uint32_t* s_cnt, float th,
                                    uint32_t octree, int max_mb, unsigned long long* stat, uint32_t* leaves,
                                    uint32_t* splits) {
#pragma clang fp contract(off)
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_leaves) return;
  bool refine = max_mb < 0 || stat[kStatFootprint] / 1000000ull < (unsigned long long)max_mb;
  uint32_t d = stree[order[q]] & ~kLeafBit;
  const uint32_t lb = octree ? 3u : 1u, kmax = octree ? 10u : 30u;
  float w = b_sw[d];
  uint32_t k = 0;
  while (refine && w > th) {
    if (k == kmax) {
      atomicOr(reinterpret_cast<unsigned int*>(stat + kStatErr), kErrStreeSize);
      break;
    }
    w = octree ? w / 8.0f : w / 2.0f;
    ++k;
  }
  uint32_t nl = 1u << (lb * k);
  leaves[q] = nl;
  splits[q] = k;
  atomicAdd(stat + kStatLeaves, (unsigned long long)nl);
  atomicAdd(stat + kStatBuildNodes, (unsigned long long)nl * reset_cnt[d]);
  atomicAdd(stat + kStatSampNodes, (unsigned long long)nl * s_cnt[d]);
}

// New node array, first the old nodes: leaf flag, source DTree and depth.
__global__ void k_stree_copy(const uint32_t* stree, uint32_t n, const uint32_t* leaf_depth, uint32_t* out,
                             uint32_t* is_leaf, uint32_t* src, uint32_t* depth) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t w = stree[i];
  out[i] = w;
  is_leaf[i] = (w & kLeafBit) ? 1u : 0u;
  if (w & kLeafBit) {
    src[i] = w & ~kLeafBit;
    depth[i] = leaf_depth[w & ~kLeafBit];
  }
}

// One work-item per leaf g of the new tree (walk order): it finds its old
// leaf q (last q with L[q] <= g) and its rank t among q's b^k leaves, and
// descends q's subtree by the base-b digits of t, most significant first:
// in a split unit placed at p (binary: the pair p, p + 1; octree: inner pairs
// p..p+5 whose children are p + 2 + 2m, leaves p + 6..p + 13) the walk visits
// the unit's leaves from the last node down, so digit u selects node
// p + U - 1 - u (U = 2 or 14), and the subtree below it occupies
// [p + U + u S, p + U + (u + 1) S) with S = 2 (b^(r-1) - 1) for r splits left.
// The first leaf of a subtree writes its split unit; every node is written once.
__global__ void k_stree_emit(const uint32_t* stree_old, uint32_t n_old, const uint32_t* order_old,
                             uint32_t n_leaves_old, const uint32_t* leaves_excl, const uint32_t* splits,
                             const uint32_t* leaf_depth, uint32_t octree, uint32_t n_leaves, uint32_t* out,
                             uint32_t* is_leaf, uint32_t* src, uint32_t* depth, uint32_t* order_new) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_leaves) return;
  uint32_t lo = 0, hi = n_leaves_old - 1u;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1u) / 2u;
    if (leaves_excl[mid] <= g) lo = mid; else hi = mid - 1u;
  }
  const uint32_t q = lo, t = g - leaves_excl[q], k = splits[q];
  uint32_t node = order_old[q];
  const uint32_t d = stree_old[node] & ~kLeafBit;
  uint32_t dep = leaf_depth[d];
  if (k > 0u) {
    const uint32_t lb = octree ? 3u : 1u, U = octree ? 14u : 2u;
    uint32_t p = n_old + 2u * (leaves_excl[q] - q);
    for (uint32_t r = k; r > 0u; --r) {
      const uint32_t shift = lb * (r - 1u);
      if ((t & ((1u << (lb * r)) - 1u)) == 0u) {
        out[node] = p;
        is_leaf[node] = 0u;
        if (octree) {
          for (uint32_t m = 0; m < 6u; ++m) {
            out[p + m] = p + 2u + 2u * m;
            is_leaf[p + m] = 0u;
          }
        }
      }
      const uint32_t u = (t >> shift) & ((1u << lb) - 1u);
      const uint32_t S = 2u * ((1u << shift) - 1u);
      node = p + U - 1u - u;
      p = p + U + u * S;
      dep += lb;
    }
  }
  is_leaf[node] = 1u;
  src[node] = d;
  depth[node] = dep;
  order_new[g] = node;
}

// Leaves numbered in node order: node word, source of each new DTree and the
// depth of its leaf.
__global__ void k_stree_finalize(uint32_t* out, uint32_t n, const uint32_t* is_leaf, const uint32_t* rank,
                                 const uint32_t* src, const uint32_t* depth, uint32_t* map, uint32_t* leaf_depth) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !is_leaf[i]) return;
  uint32_t r = rank[i];
  out[i] = kLeafBit | r;
  map[r] = src[i];
  leaf_depth[r] = depth[i];
}

// Node counts of the new DTrees from their sources.
__global__ void k_dtree_gather_counts(const uint32_t* map, uint32_t n, const uint32_t* s_cnt_old,
                                      const uint32_t* reset_cnt, uint32_t* s_cnt, uint32_t* b_cnt) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t sj = map[j];
  s_cnt[j] = s_cnt_old[sj];
  b_cnt[j] = reset_cnt[sj];
}

// Lookup table: the node word after the first 3 G levels of the descent for
// every cell of the 2^G cubed grid (child along an axis = bit G-1-level of the
// cell coordinate; the leaf word where the descent ends earlier).
__global__ void k_stree_jump(const uint32_t* stree, uint32_t G, uint32_t* jump) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = 1u << G;
  if (c >= n * n * n) return;
  uint32_t xyz[3] = {c % n, (c / n) % n, c / (n * n)};
  uint32_t w = stree[0];
  for (uint32_t k = 0; k < 3u * G && !(w & kLeafBit); ++k) {
    uint32_t bit = (xyz[k % 3u] >> (G - 1u - k / 3u)) & 1u;
    w = stree[w + bit];
  }
  jump[c] = w;
}

// DTreeT::Impl<Nearest>::RecordIrradiance (:419-427) at the DTree found by
// the STree lookup of `pos`; with v.filter set, the stochastic box filters
// (stree_box_jitter, dtree_box_jitter) move the record first, drawing five
// numbers from filter_rng(seed, i, batch, 0).
__global__ void k_sdtree_record(SdTreeView v, const float* pos, const float* canon, const float* irr,
                                const float* sw, uint32_t n, uint64_t seed, uint32_t batch) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 p{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  uint32_t cell = stree_lookup(v, p);
  float cu = canon[2 * i], cv = canon[2 * i + 1];
  bool flux = true;
  if (v.filter) {
    Pcg32 fr = filter_rng(seed, i, batch, 0u);
    float f0 = rnd(fr), f1 = rnd(fr), f2 = rnd(fr), f3 = rnd(fr), f4 = rnd(fr);
    if (v.filter & kFilterSpatialBox) cell = stree_box_jitter(v, p, cell, f0, f1, f2);
    if (cell == kInvalidCell) return;
    if (v.filter & kFilterDirectionalBox) flux = dtree_box_jitter(v, cell, cu, cv, f3, f4);
  }
  float w = sw[i];
  if (!(isfinite(w) && w > 0.0f)) return;
  unsafeAtomicAdd(v.build_sw + cell, w);
  float val = irr[i];
  if (!(isfinite(val) && val > 0.0f) || !flux) return;
  uint32_t base = fbits(v.info[cell].w);
  uint32_t node = 0;
  int idx = dtree_child_idx(cu, cv);
  for (;;) {
    uint32_t ch = dnode_child(v.build_child[base + node], idx);
    if (ch == 0) break;
    node = ch;
    idx = dtree_child_idx(cu, cv);
  }
  unsafeAtomicAdd(reinterpret_cast<float*>(v.build_sums + base + node) + idx, val * w);
}

__global__ void k_sdtree_sample(SdTreeView v, const uint32_t* cells, uint32_t n, uint64_t seed, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pcg32 rng = pcg32_seed(seed, i);
  float u, w, pdf;
  // chain b evaluates a point spread over the square beside the sampling
  // chain; a lone evaluation must give the same bits, else u is NaN
  float bx = check_coord(i, 2654435769u), by = check_coord(i, 0x85EBCA6Bu), pb, pb1, unused;
  dtree_pair(v, cells[i], true, rng, u, w, pdf, true, bx, by, pb);
  dtree_pair(v, cells[i], false, rng, bx, by, pb1, false, 0.0f, 0.0f, unused);
  if (__float_as_uint(pb) != __float_as_uint(pb1)) u = __uint_as_float(0x7FC00000u);
  out[2 * i] = u;
  out[2 * i + 1] = w;
}

__global__ void k_sdtree_pdf(SdTreeView v, const uint32_t* cells, const float* canon, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pcg32 rng = pcg32_seed(0u, i);  // drawn from by neither chain
  float u = canon[2 * i], w = canon[2 * i + 1], pa, pb;
  dtree_pair(v, cells[i], false, rng, u, w, pa, true, canon[2 * i], canon[2 * i + 1], pb);
  out[i] = __float_as_uint(pa) == __float_as_uint(pb) ? pa : __uint_as_float(0x7FC00000u);
}

__global__ void k_sdtree_lookup(SdTreeView v, const float* pos, uint32_t n, uint32_t* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = stree_lookup(v, V3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
}

// ------------------------------------------------------------------ launchers
hipError_t launch_dtree_build(DTreeArrays a, float* area_tmp, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dtree_build, dim3(nblk(a.n, 64)), dim3(64), 0, s, a, area_tmp);
  return hipGetLastError();
}
hipError_t launch_dtree_publish(DTreeArrays a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dtree_publish, dim3(nblk(a.n, 256)), dim3(256), 0, s, a);
  if (a.s_nodes && a.s_node)
    hipLaunchKernelGGL(k_dtree_interleave, dim3(nblk(a.s_nodes, 256)), dim3(256), 0, s, a);
  if (a.s_node && a.s_jump)
    hipLaunchKernelGGL(k_dtree_pdf_jump, dim3(nblk(a.n * kPdfJumpCells, 256)), dim3(256), 0, s, a);
  if (a.b_jump)
    hipLaunchKernelGGL(k_dtree_record_jump, dim3(nblk(a.n * kPdfJumpCells, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_dtree_reset(DTreeArrays src, DTreeArrays dst, const uint32_t* map, uint32_t max_depth,
                              float th, int pass, uint32_t* err, hipStream_t s) {
  if (dst.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dtree_reset, dim3(nblk(dst.n, 64)), dim3(64), 0, s, src, dst, map, max_depth, th, pass,
                     err);
  return hipGetLastError();
}
hipError_t launch_dtree_copy_sampling(DTreeArrays src, DTreeArrays dst, const uint32_t* map, hipStream_t s) {
  if (dst.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dtree_copy_sampling, dim3(dst.n), dim3(256), 0, s, src, dst, map);
  return hipGetLastError();
}
hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, hipStream_t s) {
  uint32_t ntiles = nblk(n, kScanTile);
  if (ntiles == 0) ntiles = 1;
  uint32_t* tile_sums = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&tile_sums), ntiles * sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kScanTile), 0, s, in, out, n, tile_sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1), 0, s, tile_sums, ntiles, total);
  hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(kScanTile), 0, s, out, n, tile_sums);
  e = hipGetLastError();
  hipFreeAsync(tile_sums, s);
  return e;
}
hipError_t launch_stree_footprint(const uint32_t* stree, uint32_t n, const uint32_t* b_cnt, const uint32_t* s_cnt,
                                  unsigned long long* stat, hipStream_t s) {
  uint32_t nb = std::min(nblk(n, 256), 1024u);
  hipLaunchKernelGGL(k_stree_footprint, dim3(nb), dim3(256), 0, s, stree, n, b_cnt, s_cnt, stat);
  return hipGetLastError();
}
hipError_t launch_stree_split_count(const uint32_t* stree, const uint32_t* order, uint32_t n_leaves,
                                    const float* b_sw, const uint32_t* reset_cnt, const uint32_t* s_cnt, float th,
                                    bool octree, int max_mb, unsigned long long* stat, uint32_t* leaves,
                                    uint32_t* splits, hipStream_t s) {
  hipLaunchKernelGGL(k_stree_split_count, dim3(nblk(n_leaves, 256)), dim3(256), 0, s, stree, order, n_leaves, b_sw,
                     reset_cnt, s_cnt, th, octree ? 1u : 0u, max_mb, stat, leaves, splits);
  return hipGetLastError();
}
hipError_t launch_stree_rebuild(const uint32_t* stree_old, uint32_t n_old, const uint32_t* order_old,
                                uint32_t n_leaves_old, const uint32_t* leaves_excl, const uint32_t* splits,
                                const uint32_t* leaf_depth_old, bool octree, uint32_t n_new, uint32_t n_leaves,
                                StreeScratch t, uint32_t* stree_new, uint32_t* order_new, uint32_t* map,
                                uint32_t* leaf_depth_new, hipStream_t s) {
  hipLaunchKernelGGL(k_stree_copy, dim3(nblk(n_old, 256)), dim3(256), 0, s, stree_old, n_old, leaf_depth_old,
                     stree_new, t.is_leaf, t.src, t.depth);
  hipLaunchKernelGGL(k_stree_emit, dim3(nblk(n_leaves, 256)), dim3(256), 0, s, stree_old, n_old, order_old,
                     n_leaves_old, leaves_excl, splits, leaf_depth_old, octree ? 1u : 0u, n_leaves, stree_new,
                     t.is_leaf, t.src, t.depth, order_new);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_exclusive_scan(t.is_leaf, t.rank, n_new, t.total, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stree_finalize, dim3(nblk(n_new, 256)), dim3(256), 0, s, stree_new, n_new, t.is_leaf, t.rank,
                     t.src, t.depth, map, leaf_depth_new);
  return hipGetLastError();
}
hipError_t launch_dtree_gather_counts(const uint32_t* map, uint32_t n, const uint32_t* s_cnt_old,
                                      const uint32_t* reset_cnt, uint32_t* s_cnt, uint32_t* b_cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_dtree_gather_counts, dim3(nblk(n, 256)), dim3(256), 0, s, map, n, s_cnt_old, reset_cnt, s_cnt,
                     b_cnt);
  return hipGetLastError();
}
hipError_t launch_stree_jump(const uint32_t* stree, uint32_t G, uint32_t* jump, hipStream_t s) {
  uint32_t cells = 1u << (3u * G);
  hipLaunchKernelGGL(k_stree_jump, dim3(nblk(cells, 256)), dim3(256), 0, s, stree, G, jump);
  return hipGetLastError();
}
hipError_t launch_sdtree_record(SdTreeView v, const float* pos, const float* canon, const float* irr,
                                const float* sw, uint32_t n, uint64_t seed, uint32_t batch, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sdtree_record, dim3(nblk(n, 256)), dim3(256), 0, s, v, pos, canon, irr, sw, n, seed, batch);
  return hipGetLastError();
}
hipError_t launch_sdtree_sample(SdTreeView v, const uint32_t* cells, uint32_t n, uint64_t seed, float* out,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sdtree_sample, dim3(nblk(n, 256)), dim3(256), 0, s, v, cells, n, seed, out);
  return hipGetLastError();
}
hipError_t launch_sdtree_pdf(SdTreeView v, const uint32_t* cells, const float* canon, uint32_t n, float* out,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sdtree_pdf, dim3(nblk(n, 256)), dim3(256), 0, s, v, cells, canon, n, out);
  return hipGetLastError();
}
hipError_t launch_sdtree_lookup(SdTreeView v, const float* pos, uint32_t n, uint32_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sdtree_lookup, dim3(nblk(n, 256)), dim3(256), 0, s, v, pos, n, out);
  return hipGetLastError();
}

}  // namespace rtpg
