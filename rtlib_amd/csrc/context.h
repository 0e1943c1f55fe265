// Library-side state behind the C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/rtpg.h"
#include "bvh.h"
#include "host.h"
#include "mlp_dev.h"
#include "render.h"
#include "sdtree.h"
#include "views.h"

struct rtpg_context_t {
  int device = 0;
  // SD-trees, hash grids, Morton quadtrees and learned guides created on this
  // context and not yet destroyed (they use its device and stream)
  int children = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  // scene
  rtpg::DBuf<rtpg::BvhNode4> nodes;
  rtpg::DBuf<rtpg::TriRec> tris;
  rtpg::DBuf<float4> shade, spec, lights;
  rtpg::DBuf<uint32_t> info;
  rtpg::DBuf<float> light_cdf;
  // optional texture coordinates + texture indices, vertex normals, textures
  rtpg::DBuf<float4> tattr, vnorm;
  rtpg::DBuf<uint4> tex;
  rtpg::DBuf<uint32_t> texels;
  bool has_tattr = false, has_vnorm = false;
  std::vector<uint32_t> order;       // leaf slot -> original triangle
  std::vector<uint32_t> slot_of;     // original triangle -> leaf slot
  std::vector<float> shade_host;     // leaf order, 4 per triangle
  std::vector<float4> lights_host;   // 4 per light, as on the device
  std::vector<uint32_t> light_tri;   // light -> original triangle
  // path-replay adjoint buffers, kept between calls
  rtpg::DBuf<float> adj_dldi, adj_galb, adj_gle, adj_scratch;
  rtpg::DBuf<float4> adj_ltot;
  std::vector<float> adj_galb_host;
  uint32_t n_tris = 0, n_lights = 0, n_nodes = 0;
  uint32_t bvh_depth = 0, max_stack = 0, max_leaf = 0;
  float world[6] = {0, 0, 0, 0, 0, 0};
  bool has_spec = false;
  // frame
  uint32_t width = 0, height = 0;
  rtpg::DBuf<float> accum;
  uint32_t accum_spp = 0;
  rtpg::DBuf<float> moments;  // per pixel (sum, sum of squares) of the last RTPG_FLAG_MOMENTS render
  uint32_t moments_spp = 0;
  rtpg::DBuf<double> relvar_part;  // per-work-group sums of rtpg_render_relvar
  rtpg::Camera cam{};
  // path arena
  rtpg::DBuf<float4> p_ray_o, p_ray_d, p_T, p_L, p_hit, p_shadow, p_vA, p_vB, p_vC;
  rtpg::DBuf<float4> p_ray_o1, p_ray_d1, p_T1;  // second slot buffers of the path state
  rtpg::DBuf<float4> p_mix, p_mv;  // learned guide: mixture per queue slot; training vertices (5 arrays)
  rtpg::DBuf<float> p_mixc;
  rtpg::DBuf<uint2> p_rng, p_rng1;
  rtpg::DBuf<uint32_t> p_nhit, p_q0, p_q1, p_count, p_deep, p_err;
  rtpg::DBuf<unsigned long long> p_stats;
  uint64_t last_stats[3] = {0, 0, 0};
  // profiling: bit 0 times kernel classes with events, bit 1 counts traversal work
  uint32_t prof_mode = 0;
  std::vector<hipEvent_t> events;
  double kernel_ms[rtpg::kTimedKernels] = {0, 0, 0, 0, 0};
  uint64_t kernel_launches[rtpg::kTimedKernels] = {0, 0, 0, 0, 0};
  rtpg::DBuf<unsigned long long> p_work;
  uint64_t work[4] = {0, 0, 0, 0};
  ~rtpg_context_t() {
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }

  rtpg::SceneView scene_view() const;
};

struct rtpg_sdtree_t {
  rtpg_context_t* ctx = nullptr;
  float aabb_min[3] = {0, 0, 0};
  float size = 1.0f;
  uint32_t max_dtree_depth = 20;
  float fraction = 0.3f;
  int split_mode = 0;
  float split_c = 4000.0f;
  int max_mb = 2000;
  float subdiv_th = 0.01f;
  bool octree = false;  // refine splits a leaf into 8 (RTSTree2T) instead of 2
  // Spatial tree on the device: node words (leaf: kLeafBit | DTree, inner:
  // index of the first of its two children), the leaves in the refine's walk
  // order, the depth of every DTree's leaf, and the lookup table of the first
  // 3 * kJumpBits levels; the *2 buffers are the other half of the refine's
  // ping-pong.
  uint32_t n_snodes = 0;
  rtpg::DBuf<uint32_t> d_stree, d_stree2;
  rtpg::DBuf<uint32_t> d_order, d_order2;
  rtpg::DBuf<uint32_t> d_leaf_depth, d_leaf_depth2;
  rtpg::DBuf<uint32_t> d_jump;
  static constexpr uint32_t kJumpBits = 5;
  // recording filters (RecordFilter bits); rtpg_sdtree_record draws from
  // filter_rng(filter_seed, i, record_batch, 0), record_batch counting its calls
  uint32_t filter = 0;
  uint64_t filter_seed = 0;
  uint32_t record_batch = 0;
  // host-side topology (import): node words, walk order and leaf depths uploaded, jump table built
  void upload_stree(const std::vector<uint32_t>& words);
  // device DTree arrays (owned)
  struct Owned {
    uint32_t n = 0, b_nodes = 0, s_nodes = 0;
    rtpg::DBuf<uint32_t> b_base, b_cnt, s_base, s_cnt, s_depth, b_depth;
    rtpg::DBuf<float> b_sw, b_sum, b_area, s_sw, s_sum, s_area;
    rtpg::DBuf<float4> info, b_sums, s_sums;
    rtpg::DBuf<uint2> b_child, s_child;
    rtpg::DBuf<uint4> s_node;  // interleaved sampling nodes (written by publish)
    rtpg::DBuf<uint4> s_jump;  // pdf jump table (written by publish)
    rtpg::DBuf<uint2> b_jump;  // record jump table (written by publish)
    rtpg::DTreeArrays view();
  } dt, spare;  // spare: the other set of reset()'s ping-pong
  rtpg::DBuf<float> area_tmp;
  rtpg::DBuf<uint32_t> d_map, d_total;
  // refine temporaries: per old DTree (reset counts, leaves, splits, their
  // prefix sums), per new node (leaf flag, rank, source, depth), counters
  rtpg::DBuf<uint32_t> d_cnt_old, d_leaves, d_splits, d_leaves_excl;
  rtpg::DBuf<uint32_t> d_is_leaf, d_rank, d_src, d_depth;
  rtpg::DBuf<unsigned long long> d_stat;

  rtpg::SdTreeView view();
  void clear();
  void build();
  void reset(int iter, int spp);
  void publish();
};

// Learned directional guide (no counterpart in the reference).
struct rtpg_mlp_t {
  rtpg_context_t* ctx = nullptr;
  rtpg_mlp_config cfg{};
  float lo[3] = {0, 0, 0};
  float size = 1.0f;
  float res[rtpg::kMlpLevels] = {};
  uint32_t dense = 0;
  size_t entries = 0;  // per level
  rtpg::DBuf<float> w, w_m1, w_m2, w_grad;
  rtpg::DBuf<uint16_t> w_bf16;
  rtpg::DBuf<float> t, t_m1, t_m2, t_grad;
  rtpg::DBuf<uint32_t> t_grad_pk;  // packed bf16 gradient of levels kMlpPkFirst.. (entries each)
  rtpg::DBuf<uint32_t> t_bf16;
  rtpg::DBuf<float> partial, loss_partial, loss;
  rtpg::DBuf<float4> samples;  // recorded by renders: capacity x 4
  rtpg::DBuf<uint32_t> count;
  rtpg::DBuf<float4> ordered;  // training windows in spatial order
  rtpg::DBuf<uint32_t> sort_hist, sort_offs, sort_total;
  rtpg::DBuf<float4> staging;
  rtpg::DBuf<float> out;
  uint64_t steps = 0;
  rtpg::DBuf<float4> cache_mix;  // spatial cache (cfg.cache_res > 0)
  rtpg::DBuf<float> cache_c;
  bool cache_stale = true;       // the network changed since the last fill
  size_t table_floats() const { return (size_t)rtpg::kMlpLevels * entries * rtpg::kMlpFeat; }
  size_t pk_entries() const { return (size_t)(rtpg::kMlpLevels - rtpg::kMlpPkFirst) * entries; }
  uint32_t cache_entries() const { return cfg.cache_res * cfg.cache_res * cfg.cache_res * rtpg::kMlpCacheFaces; }
};

rtpg::MlpView mlp_view_of(const rtpg_mlp_t* m, bool training);
// Refills the spatial cache if the network changed since the last fill.
void mlp_cache_refresh(rtpg_mlp_t* m);
