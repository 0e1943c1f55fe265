// SD-tree host logic and C-ABI: spatial refine (host, topology only), device
// DTree reset/build driving, export/import and the pass controller.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rtpg.h"
#include "context.h"

using namespace rtpg;

namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    return RTPG_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RTPG_ERR_INVALID;
  } catch (...) {
    set_last_error("unknown error");
    return RTPG_ERR_INVALID;
  }
}

void require(bool c, const char* msg) {
  if (!c) throw Error(RTPG_ERR_INVALID, msg);
}

std::vector<uint32_t> exclusive_prefix(const std::vector<uint32_t>& c, uint32_t& total) {
  std::vector<uint32_t> b(c.size());
  uint64_t acc = 0;
  for (size_t i = 0; i < c.size(); ++i) {
    b[i] = (uint32_t)acc;
    acc += c[i];
  }
  if (acc > 0xFFFFFFFFull) throw Error(RTPG_ERR_LIMIT, "DTree node pool exceeds 2^32 nodes");
  total = (uint32_t)acc;
  return b;
}

}  // namespace

DTreeArrays rtpg_sdtree_t::Owned::view() {
  DTreeArrays a;
  a.n = n;
  a.b_nodes = b_nodes;
  a.s_nodes = s_nodes;
  a.b_base = b_base.p;
  a.b_cnt = b_cnt.p;
  a.s_base = s_base.p;
  a.s_cnt = s_cnt.p;
  a.s_depth = s_depth.p;
  a.b_depth = b_depth.p;
  a.b_sw = b_sw.p;
  a.b_sum = b_sum.p;
  a.b_area = b_area.p;
  a.s_sw = s_sw.p;
  a.s_sum = s_sum.p;
  a.s_area = s_area.p;
  a.info = info.p;
  a.b_sums = b_sums.p;
  a.b_child = b_child.p;
  a.s_sums = s_sums.p;
  a.s_child = s_child.p;
  a.s_node = s_node.p;
  a.s_jump = s_jump.p;
  a.b_jump = b_jump.p;
  return a;
}

// Per-DTree arrays of at least n entries (buffers only grow, so the ping-pong
// sets of reset() stop allocating once the tree has reached its size).
static void alloc_per_dtree(rtpg_sdtree_t::Owned& o, uint32_t n) {
  o.n = n;
  o.b_base.reserve(n);
  o.b_cnt.reserve(n);
  o.s_base.reserve(n);
  o.s_cnt.reserve(n);
  o.s_depth.reserve(n);
  o.b_depth.reserve(n);
  o.b_sw.reserve(n);
  o.b_sum.reserve(n);
  o.b_area.reserve(n);
  o.s_sw.reserve(n);
  o.s_sum.reserve(n);
  o.s_area.reserve(n);
  o.info.reserve(n);
}

// Node words of a validated topology to the device, with the leaves in the
// refine's walk order (depth first, second child first) and the depth of
// every DTree's leaf (root = 0, for the voxel size of the spatial box filter);
// the lookup table is built on the device.
void rtpg_sdtree_t::upload_stree(const std::vector<uint32_t>& words) {
  hipStream_t s = ctx->stream;
  std::vector<uint32_t> order, depth;
  std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 0u}};  // (node, depth)
  while (!stack.empty()) {
    auto [i, d] = stack.back();
    stack.pop_back();
    uint32_t w = words[i];
    if (w & kLeafBit) {
      uint32_t t = w & ~kLeafBit;
      if (t >= depth.size()) depth.resize((size_t)t + 1u, 0u);
      depth[t] = d;
      order.push_back(i);
    } else {
      stack.push_back({w, d + 1u});
      stack.push_back({w + 1u, d + 1u});
    }
  }
  n_snodes = (uint32_t)words.size();
  d_stree.upload(words, s);
  d_order.upload(order, s);
  d_leaf_depth.upload(depth, s);
  d_jump.reserve((size_t)1 << (3 * kJumpBits));
  RTPG_HIP(launch_stree_jump(d_stree.p, kJumpBits, d_jump.p, s));
}

SdTreeView rtpg_sdtree_t::view() {
  SdTreeView v{};
  v.stree = d_stree.p;
  v.jump = d_jump.p;
  v.jump_bits = kJumpBits;
  for (int k = 0; k < 3; ++k) v.aabb_min[k] = aabb_min[k];
  v.aabb_size = size;
  v.info = dt.info.p;
  v.build_sw = dt.b_sw.p;
  v.build_sums = dt.b_sums.p;
  v.build_child = dt.b_child.p;
  v.samp_sums = dt.s_sums.p;
  v.samp_child = dt.s_child.p;
  v.samp_node = dt.s_node.p;
  v.samp_jump = dt.s_jump.p;
  v.build_jump = dt.b_jump.p;
  v.fraction = fraction;
  v.n_dtrees = dt.n;
  v.filter = filter;
  v.leaf_depth = d_leaf_depth.p;
  return v;
}

namespace rtpg {
SdTreeView sdtree_view_of(rtpg_sdtree_t* t) { return t->view(); }
}  // namespace rtpg

void rtpg_sdtree_t::publish() {
  dt.s_node.reserve(2 * (size_t)std::max(dt.s_nodes, 1u));
  dt.s_jump.reserve((size_t)kPdfJumpCells * std::max(dt.n, 1u));
  dt.b_jump.reserve((size_t)kPdfJumpCells * std::max(dt.n, 1u));
  RTPG_HIP(launch_dtree_publish(dt.view(), ctx->stream));
}

// RTSTreeWrapperT::Clear (OPX7UtilsPathGuiding.h:1634-1636): a single leaf
// whose DTrees (building, sampling) are one empty node each (RTDTreeT ctor :879-886).
void rtpg_sdtree_t::clear() {
  hipStream_t s = ctx->stream;
  upload_stree({kLeafBit});
  alloc_per_dtree(dt, 1);
  std::vector<uint32_t> zero_u{0u}, one_u{1u};
  std::vector<float> zero_f{0.0f};
  dt.b_base.upload(zero_u, s);
  dt.s_base.upload(zero_u, s);
  dt.b_cnt.upload(one_u, s);
  dt.s_cnt.upload(one_u, s);
  dt.b_depth.upload(zero_u, s);
  dt.s_depth.upload(zero_u, s);
  for (auto* f : {&dt.b_sw, &dt.b_sum, &dt.b_area, &dt.s_sw, &dt.s_sum, &dt.s_area}) f->upload(zero_f, s);
  dt.b_nodes = dt.s_nodes = 1;
  dt.b_sums.alloc(1);
  dt.s_sums.alloc(1);
  dt.b_child.alloc(1);
  dt.s_child.alloc(1);
  dt.b_sums.zero(1, s);
  dt.s_sums.zero(1, s);
  dt.b_child.zero(1, s);
  dt.s_child.zero(1, s);
  publish();
  RTPG_HIP(hipStreamSynchronize(s));
}

// RTSTreeWrapperT::Build (:1657-1667): per leaf building.Build(), then
// sampling = building.
void rtpg_sdtree_t::build() {
  hipStream_t s = ctx->stream;
  area_tmp.reserve(dt.b_nodes);
  RTPG_HIP(launch_dtree_build(dt.view(), area_tmp.p, s));
  uint32_t n = dt.n;
  dt.s_sums.reserve(dt.b_nodes);
  dt.s_child.reserve(dt.b_nodes);
  dt.s_nodes = dt.b_nodes;
  RTPG_HIP(hipMemcpyAsync(dt.s_sums.p, dt.b_sums.p, sizeof(float4) * dt.b_nodes, hipMemcpyDeviceToDevice, s));
  RTPG_HIP(hipMemcpyAsync(dt.s_child.p, dt.b_child.p, sizeof(uint2) * dt.b_nodes, hipMemcpyDeviceToDevice, s));
  auto cp_u = [&](DBuf<uint32_t>& d, DBuf<uint32_t>& src) {
    RTPG_HIP(hipMemcpyAsync(d.p, src.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  };
  auto cp_f = [&](DBuf<float>& d, DBuf<float>& src) {
    RTPG_HIP(hipMemcpyAsync(d.p, src.p, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
  };
  cp_u(dt.s_base, dt.b_base);
  cp_u(dt.s_cnt, dt.b_cnt);
  cp_u(dt.s_depth, dt.b_depth);
  cp_f(dt.s_sw, dt.b_sw);
  cp_f(dt.s_sum, dt.b_sum);
  cp_f(dt.s_area, dt.b_area);
  // no host wait: what follows (a render, the next reset's downloads) is
  // ordered on the same stream
  publish();
}

// RTSTreeWrapperT::Reset (:1645-1656): refine (RTSTreeT::Refine :1275-1311)
// then per leaf RTDTreeWrapperT::Reset(maxDepth, 0.01), all on the device; the
// host reads five counters once to size the new arrays.
void rtpg_sdtree_t::reset(int iter, int spp) {
  if (iter <= 0) return;
  hipStream_t s = ctx->stream;
  size_t th;
  if (split_mode == 0) {
    th = (size_t)(std::sqrt(std::pow(2.0, iter) * spp / 4.0f) * (double)split_c);
  } else {
    th = (size_t)((double)split_c * std::sqrt(std::pow(2.0, iter)));
  }
  const uint32_t n = dt.n, nn = n_snodes;
  // 1. reset node counts of every old DTree, split counts of the old leaves
  //    in walk order and their prefix sums
  d_stat.reserve(kStatWords);
  d_stat.zero(kStatWords, s);
  uint32_t* err = reinterpret_cast<uint32_t*>(d_stat.p + kStatErr);
  d_cnt_old.reserve(n);
  d_leaves.reserve(n);
  d_splits.reserve(n);
  d_leaves_excl.reserve(n);
  d_total.reserve(1);
  DTreeArrays src = dt.view();
  DTreeArrays cnt_view = src;
  cnt_view.b_cnt = d_cnt_old.p;
  RTPG_HIP(launch_dtree_reset(src, cnt_view, nullptr, max_dtree_depth, subdiv_th, 0, err, s));
  if (max_mb >= 0) RTPG_HIP(launch_stree_footprint(d_stree.p, nn, dt.b_cnt.p, dt.s_cnt.p, d_stat.p, s));
  RTPG_HIP(launch_stree_split_count(d_stree.p, d_order.p, n, dt.b_sw.p, d_cnt_old.p, dt.s_cnt.p, (float)th, octree,
                                    max_mb, d_stat.p, d_leaves.p, d_splits.p, s));
  RTPG_HIP(launch_exclusive_scan(d_leaves.p, d_leaves_excl.p, n, d_total.p, s));
  unsigned long long stat[kStatWords];
  d_stat.download(stat, kStatWords, s);  // the one host wait of the refine
  if (stat[kStatErr] & kErrResetStack) throw Error(RTPG_ERR_LIMIT, "DTree reset stack bound exceeded");
  const unsigned long long n2 = stat[kStatLeaves];
  const unsigned long long nn2 = (unsigned long long)nn + 2ull * (n2 - n);
  if ((stat[kStatErr] & kErrStreeSize) || nn2 >= 0x80000000ull)
    throw Error(RTPG_ERR_LIMIT, "spatial tree exceeds 2^31 nodes");
  if (stat[kStatBuildNodes] > 0xFFFFFFFFull || stat[kStatSampNodes] > 0xFFFFFFFFull)
    throw Error(RTPG_ERR_LIMIT, "DTree nodes exceed 2^32");
  // 2. new node array, walk order, DTree sources and leaf depths
  d_stree2.reserve(nn2);
  d_order2.reserve(n2);
  d_leaf_depth2.reserve(n2);
  d_map.reserve(n2);
  d_is_leaf.reserve(nn2);
  d_rank.reserve(nn2);
  d_src.reserve(nn2);
  d_depth.reserve(nn2);
  StreeScratch t{d_is_leaf.p, d_rank.p, d_src.p, d_depth.p, d_total.p};
  RTPG_HIP(launch_stree_rebuild(d_stree.p, nn, d_order.p, n, d_leaves_excl.p, d_splits.p, d_leaf_depth.p, octree,
                                (uint32_t)nn2, (uint32_t)n2, t, d_stree2.p, d_order2.p, d_map.p, d_leaf_depth2.p, s));
  // 3. new DTree arrays in the spare set (swapped in at the end): sampling
  //    trees copied from the sources, building trees reset from them
  Owned& nd = spare;
  alloc_per_dtree(nd, (uint32_t)n2);
  RTPG_HIP(launch_dtree_gather_counts(d_map.p, (uint32_t)n2, dt.s_cnt.p, d_cnt_old.p, nd.s_cnt.p, nd.b_cnt.p, s));
  RTPG_HIP(launch_exclusive_scan(nd.s_cnt.p, nd.s_base.p, (uint32_t)n2, d_total.p, s));
  RTPG_HIP(launch_exclusive_scan(nd.b_cnt.p, nd.b_base.p, (uint32_t)n2, d_total.p, s));
  nd.s_nodes = (uint32_t)stat[kStatSampNodes];
  nd.b_nodes = (uint32_t)stat[kStatBuildNodes];
  nd.s_sums.reserve(nd.s_nodes);
  nd.s_child.reserve(nd.s_nodes);
  nd.b_sums.reserve(nd.b_nodes);
  nd.b_child.reserve(nd.b_nodes);
  DTreeArrays dstv = nd.view();
  RTPG_HIP(launch_dtree_copy_sampling(src, dstv, d_map.p, s));
  RTPG_HIP(launch_dtree_reset(src, dstv, d_map.p, max_dtree_depth, subdiv_th, 1, err, s));
  RTPG_HIP(launch_stree_jump(d_stree2.p, kJumpBits, d_jump.p, s));
  std::swap(d_stree, d_stree2);
  std::swap(d_order, d_order2);
  std::swap(d_leaf_depth, d_leaf_depth2);
  n_snodes = (uint32_t)nn2;
  std::swap(dt, spare);
  publish();
}

// ------------------------------------------------------------------ C ABI
extern "C" {

int rtpg_sdtree_create(rtpg_context ctx, const float aabb_min[3], const float aabb_max[3], uint32_t max_dtree_depth,
                       float fraction, rtpg_sdtree* out) {
  return guarded([&] {
    require(ctx && aabb_min && aabb_max && out, "NULL argument");
    require(max_dtree_depth >= 1 && max_dtree_depth <= 20, "max_dtree_depth must be in [1, 20]");
    require(fraction >= 0.0f && fraction <= 1.0f, "fraction must be in [0, 1]");
    RTPG_HIP(hipSetDevice(ctx->device));
    auto* t = new rtpg_sdtree_t();
    t->ctx = ctx;
    // RTSTreeT ctor (:1202-1208): cube of the largest extent at aabbMin.
    float sx = aabb_max[0] - aabb_min[0], sy = aabb_max[1] - aabb_min[1], sz = aabb_max[2] - aabb_min[2];
    float m = std::max(std::max(sx, sy), sz);
    for (int k = 0; k < 3; ++k) t->aabb_min[k] = aabb_min[k];
    // size = (aabbMin + maxSize) - aabbMin, as the device computes aabbMax - aabbMin
    float mx = aabb_min[0] + m;
    t->size = mx - aabb_min[0];
    t->max_dtree_depth = max_dtree_depth;
    t->fraction = fraction;
    try {
      t->clear();
    } catch (...) {
      delete t;
      throw;
    }
    t->ctx->children++;
    *out = t;
  });
}

int rtpg_sdtree_destroy(rtpg_sdtree tree) {
  return guarded([&] {
    if (!tree) return;
    (void)hipStreamSynchronize(tree->ctx->stream);
    tree->ctx->children--;
    delete tree;
  });
}

int rtpg_sdtree_clear(rtpg_sdtree tree) {
  return guarded([&] {
    require(tree != nullptr, "tree is NULL");
    tree->clear();
  });
}

int rtpg_sdtree_set_split(rtpg_sdtree tree, int mode, float c, int max_mb, float subdiv_th) {
  return guarded([&] {
    require(tree != nullptr, "tree is NULL");
    require(mode == 0 || mode == 1, "mode must be 0 or 1");
    require(c > 0.0f, "c must be positive");
    // Each subdivided slot holds more than subdiv_th of the mass, so a level
    // has at most 1/subdiv_th of them: keep the DTree under 65536 nodes.
    require(subdiv_th > 0.0f && 1.0 + 4.0 * (tree->max_dtree_depth - 1) * std::floor(1.0 / subdiv_th) < 65536.0,
            "subdiv_th too small for 16-bit DTree children");
    tree->split_mode = mode;
    tree->split_c = c;
    tree->max_mb = max_mb;
    tree->subdiv_th = subdiv_th;
  });
}

int rtpg_sdtree_reset(rtpg_sdtree tree, int iter, int samples_per_pass) {
  return guarded([&] {
    require(tree != nullptr, "tree is NULL");
    tree->reset(iter, samples_per_pass);
  });
}

int rtpg_sdtree_build(rtpg_sdtree tree) {
  return guarded([&] {
    require(tree != nullptr, "tree is NULL");
    tree->build();
  });
}

int rtpg_sdtree_record(rtpg_sdtree tree, const float* pos, const float* canon, const float* irradiance,
                       const float* stat_weight, uint32_t n) {
  return guarded([&] {
    require(tree && pos && canon && irradiance && stat_weight, "NULL argument");
    if (n == 0) return;
    hipStream_t s = tree->ctx->stream;
    DBuf<float> dp, dc, di, dw;
    dp.upload(pos, 3 * (size_t)n, s);
    dc.upload(canon, 2 * (size_t)n, s);
    di.upload(irradiance, n, s);
    dw.upload(stat_weight, n, s);
    RTPG_HIP(launch_sdtree_record(tree->view(), dp.p, dc.p, di.p, dw.p, n, tree->filter_seed, tree->record_batch++,
                                  s));
    RTPG_HIP(hipStreamSynchronize(s));
  });
}

int rtpg_sdtree_set_octree(rtpg_sdtree tree, int enable) {
  return guarded([&] {
    require(tree != nullptr, "tree is NULL");
    require(enable == 0 || enable == 1, "enable is 0 or 1");
    tree->octree = enable != 0;
  });
}

int rtpg_sdtree_set_filter(rtpg_sdtree tree, int spatial, int directional, uint64_t seed) {
  return guarded([&] {
    require(tree != nullptr, "tree is NULL");
    require((spatial == 0 || spatial == 1) && (directional == 0 || directional == 1),
            "filters are 0 (nearest) or 1 (box)");
    tree->filter = (spatial ? kFilterSpatialBox : 0u) | (directional ? kFilterDirectionalBox : 0u);
    tree->filter_seed = seed;
    tree->record_batch = 0;
  });
}

int rtpg_sdtree_lookup(rtpg_sdtree tree, const float* pos, uint32_t n, uint32_t* dtree) {
  return guarded([&] {
    require(tree && pos && dtree, "NULL argument");
    if (n == 0) return;
    hipStream_t s = tree->ctx->stream;
    DBuf<float> dp;
    DBuf<uint32_t> dout;
    dp.upload(pos, 3 * (size_t)n, s);
    dout.alloc(n);
    RTPG_HIP(launch_sdtree_lookup(tree->view(), dp.p, n, dout.p, s));
    dout.download(dtree, n, s);
  });
}

int rtpg_sdtree_sample(rtpg_sdtree tree, const uint32_t* dtree, uint32_t n, uint64_t seed, float* canon) {
  return guarded([&] {
    require(tree && dtree && canon, "NULL argument");
    if (n == 0) return;
    for (uint32_t i = 0; i < n; ++i) require(dtree[i] < tree->dt.n, "DTree index out of range");
    hipStream_t s = tree->ctx->stream;
    DBuf<uint32_t> dc;
    DBuf<float> dout;
    dc.upload(dtree, n, s);
    dout.alloc(2 * (size_t)n);
    RTPG_HIP(launch_sdtree_sample(tree->view(), dc.p, n, seed, dout.p, s));
    dout.download(canon, 2 * (size_t)n, s);
  });
}

int rtpg_sdtree_pdf(rtpg_sdtree tree, const uint32_t* dtree, const float* canon, uint32_t n, float* pdf) {
  return guarded([&] {
    require(tree && dtree && canon && pdf, "NULL argument");
    if (n == 0) return;
    for (uint32_t i = 0; i < n; ++i) require(dtree[i] < tree->dt.n, "DTree index out of range");
    hipStream_t s = tree->ctx->stream;
    DBuf<uint32_t> dc;
    DBuf<float> dq, dout;
    dc.upload(dtree, n, s);
    dq.upload(canon, 2 * (size_t)n, s);
    dout.alloc(n);
    RTPG_HIP(launch_sdtree_pdf(tree->view(), dc.p, dq.p, n, dout.p, s));
    dout.download(pdf, n, s);
  });
}

int rtpg_sdtree_sizes(rtpg_sdtree tree, uint32_t sizes[4]) {
  return guarded([&] {
    require(tree && sizes, "NULL argument");
    sizes[0] = tree->n_snodes;
    sizes[1] = tree->dt.n;
    sizes[2] = tree->dt.b_nodes;
    sizes[3] = tree->dt.s_nodes;
  });
}

int rtpg_sdtree_export(rtpg_sdtree tree, uint32_t* stree, uint32_t* b_cnt, uint32_t* s_cnt, float* scal,
                       float* b_sums, uint16_t* b_child, float* s_sums, uint16_t* s_child) {
  return guarded([&] {
    require(tree != nullptr, "tree is NULL");
    hipStream_t s = tree->ctx->stream;
    auto& dt = tree->dt;
    uint32_t n = dt.n;
    if (stree) tree->d_stree.download(stree, tree->n_snodes, s);
    if (b_cnt) dt.b_cnt.download(b_cnt, n, s);
    if (s_cnt) dt.s_cnt.download(s_cnt, n, s);
    if (scal) {
      DBuf<float>* arrs[6] = {&dt.b_sw, &dt.b_sum, &dt.b_area, &dt.s_sw, &dt.s_sum, &dt.s_area};
      for (int k = 0; k < 6; ++k) arrs[k]->download(scal + (size_t)k * n, n, s);
    }
    if (b_sums) dt.b_sums.download(reinterpret_cast<float4*>(b_sums), dt.b_nodes, s);
    if (b_child) dt.b_child.download(reinterpret_cast<uint2*>(b_child), dt.b_nodes, s);
    if (s_sums) dt.s_sums.download(reinterpret_cast<float4*>(s_sums), dt.s_nodes, s);
    if (s_child) dt.s_child.download(reinterpret_cast<uint2*>(s_child), dt.s_nodes, s);
  });
}

int rtpg_sdtree_import(rtpg_sdtree tree, uint32_t n_stree, const uint32_t* stree, uint32_t n_dtrees,
                       const uint32_t* b_cnt, const uint32_t* s_cnt, const float* scal, const float* b_sums,
                       const uint16_t* b_child, const float* s_sums, const uint16_t* s_child) {
  return guarded([&] {
    require(tree && stree && b_cnt && s_cnt && scal && b_sums && b_child && s_sums && s_child, "NULL argument");
    require(n_stree >= 1 && n_dtrees >= 1, "empty tree");
    // Validate the topology: pairs of children, every leaf names a distinct DTree.
    std::vector<uint8_t> seen(n_dtrees, 0);
    uint32_t leaves = 0;
    for (uint32_t i = 0; i < n_stree; ++i) {
      uint32_t w = stree[i];
      if (w & 0x80000000u) {
        uint32_t d = w & 0x7FFFFFFFu;
        require(d < n_dtrees && !seen[d], "bad DTree index in stree");
        seen[d] = 1;
        ++leaves;
      } else {
        require(w > i && w + 1u < n_stree, "bad child index in stree");
      }
    }
    require(leaves == n_dtrees, "stree leaves must match n_dtrees");
    std::vector<uint32_t> bc(b_cnt, b_cnt + n_dtrees), scv(s_cnt, s_cnt + n_dtrees);
    for (uint32_t j = 0; j < n_dtrees; ++j) require(bc[j] >= 1 && bc[j] <= 65536u && scv[j] >= 1 && scv[j] <= 65536u,
                                                    "DTree node counts must be in [1, 65536]");
    uint32_t bt = 0, st = 0;
    std::vector<uint32_t> bb = exclusive_prefix(bc, bt), sb = exclusive_prefix(scv, st);
    // child indices must stay inside their DTree and point forward
    for (uint32_t j = 0; j < n_dtrees; ++j) {
      for (uint32_t k = 0; k < bc[j]; ++k)
        for (int c = 0; c < 4; ++c) {
          uint32_t ch = b_child[4 * ((size_t)bb[j] + k) + c];
          require(ch == 0 || (ch > k && ch < bc[j]), "bad building child index");
        }
      for (uint32_t k = 0; k < scv[j]; ++k)
        for (int c = 0; c < 4; ++c) {
          uint32_t ch = s_child[4 * ((size_t)sb[j] + k) + c];
          require(ch == 0 || (ch > k && ch < scv[j]), "bad sampling child index");
        }
    }
    hipStream_t s = tree->ctx->stream;
    rtpg_sdtree_t::Owned o;
    alloc_per_dtree(o, n_dtrees);
    o.b_nodes = bt;
    o.s_nodes = st;
    o.b_cnt.upload(bc, s);
    o.s_cnt.upload(scv, s);
    o.b_base.upload(bb, s);
    o.s_base.upload(sb, s);
    std::vector<uint32_t> z(n_dtrees, 0u);
    o.b_depth.upload(z, s);
    o.s_depth.upload(z, s);
    DBuf<float>* arrs[6] = {&o.b_sw, &o.b_sum, &o.b_area, &o.s_sw, &o.s_sum, &o.s_area};
    for (int k = 0; k < 6; ++k) arrs[k]->upload(scal + (size_t)k * n_dtrees, n_dtrees, s);
    o.b_sums.upload(reinterpret_cast<const float4*>(b_sums), bt, s);
    o.b_child.upload(reinterpret_cast<const uint2*>(b_child), bt, s);
    o.s_sums.upload(reinterpret_cast<const float4*>(s_sums), st, s);
    o.s_child.upload(reinterpret_cast<const uint2*>(s_child), st, s);
    RTPG_HIP(hipStreamSynchronize(s));
    tree->upload_stree(std::vector<uint32_t>(stree, stree + n_stree));
    tree->dt = std::move(o);
    tree->publish();
    RTPG_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"

// ------------------------------------------------------------------ controller
// Pass schedule of RTSTreeControllerT (OPX7UtilsPathGuiding.h:1769-1920):
// passes of 2^k * spl samples; the last pass takes the remainder when what is
// left after it would be less than twice its size, or when the samples spent
// reach ratio * budget.
struct rtpg_sdtree_ctrl_t {
  rtpg_sdtree_t* tree = nullptr;
  uint32_t budget = 0, spl = 1, iter_built = 0;
  float ratio = 0.5f;
  bool start = false, executing = false;
  uint32_t per_all = 0, per_tmp = 0, remain = 0, pass = 0, iter = 0;
  uint32_t state = 0;
};

extern "C" {

int rtpg_sdtree_ctrl_create(rtpg_sdtree tree, uint32_t sample_for_budget, uint32_t iteration_for_built,
                            float ratio_for_budget, uint32_t sample_per_launch, rtpg_sdtree_ctrl* out) {
  return guarded([&] {
    require(tree && out, "NULL argument");
    require(sample_per_launch >= 1 && sample_for_budget >= 1, "budget and samples per launch must be >= 1");
    auto* c = new rtpg_sdtree_ctrl_t();
    c->tree = tree;
    c->budget = sample_for_budget;
    c->iter_built = iteration_for_built;
    c->ratio = ratio_for_budget;
    c->spl = sample_per_launch;
    *out = c;
  });
}

int rtpg_sdtree_ctrl_destroy(rtpg_sdtree_ctrl c) {
  delete c;
  return RTPG_OK;
}

int rtpg_sdtree_ctrl_start(rtpg_sdtree_ctrl c) {
  return guarded([&] {
    require(c != nullptr, "ctrl is NULL");
    c->start = true;
  });
}

int rtpg_sdtree_ctrl_beg_trace(rtpg_sdtree_ctrl c) {
  return guarded([&] {
    require(c != nullptr, "ctrl is NULL");
    if (c->start) {
      c->per_all = 0;
      c->per_tmp = 0;
      c->iter = 0;
      c->start = false;
      c->executing = true;
      c->remain = ((c->budget - 1 + c->spl) / c->spl) * c->spl;
      c->pass = 0;
      c->state = 0;
      c->tree->clear();
    }
    if (!c->executing) return;
    if (c->per_tmp == 0) {
      c->remain -= c->pass;
      uint64_t want = (uint64_t)(1ull << std::min(c->iter, 31u)) * c->spl;
      c->pass = (uint32_t)std::min<uint64_t>(c->remain, want);
      if ((c->remain - c->pass < 2 * c->pass) || ((float)c->per_all >= c->ratio * (float)c->budget)) {
        c->pass = c->remain;
      }
      if (c->remain > c->pass) c->tree->reset((int)c->iter, (int)c->pass);
    }
    if (c->iter > c->iter_built) {
      c->state = c->remain > c->pass ? 1u : 2u;
    } else {
      c->state = 0u;
    }
  });
}

int rtpg_sdtree_ctrl_end_trace(rtpg_sdtree_ctrl c, uint32_t samples) {
  return guarded([&] {
    require(c != nullptr, "ctrl is NULL");
    if (!c->executing) return;
    uint32_t k = samples ? samples : c->spl;
    c->per_all += k;
    c->per_tmp += k;
    if (c->per_tmp >= c->pass) {
      c->tree->build();
      c->iter++;
      c->per_tmp = 0;
    }
    if (c->per_all > c->budget) {
      c->executing = false;
      c->per_all = 0;
    }
  });
}

int rtpg_sdtree_ctrl_query(rtpg_sdtree_ctrl c, uint32_t q[6]) {
  return guarded([&] {
    require(c && q, "NULL argument");
    q[0] = c->state;
    q[1] = c->iter;
    q[2] = c->per_tmp;
    q[3] = c->pass;
    q[4] = c->executing ? 1u : 0u;
    q[5] = c->per_all;
  });
}

}  // extern "C"
