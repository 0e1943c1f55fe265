// SD-tree (Muller et al. 2017) as flattened device arrays.
//
// Device layout (all in HBM, structure of arrays):
//   stree[node]            u32: leaf -> 0x80000000 | DTree index, inner -> index of the
//                          first child (children are allocated in pairs, axis = depth % 3)
//   per DTree j            b_base/b_cnt, s_base/s_cnt (u32), b_sw (recorded weight),
//                          b_sum/b_area, s_sw/s_sum/s_area (f32), info (float4 view for
//                          the tracer: s_base, s_sum, s_sw, b_base)
//   node pools             sums (float4) and children (uint2 = 4 x u16, relative to the
//                          DTree's base) for the building and the sampling trees
// DTree nodes are numbered exactly as the reference numbers them (DFS of
// RTDTreeT::Reset), so exported arrays compare element-wise with it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "views.h"

namespace rtpg {

struct DTreeArrays {
  uint32_t n = 0;          // DTrees
  uint32_t b_nodes = 0;    // building pool size
  uint32_t s_nodes = 0;    // sampling pool size
  uint32_t* b_base = nullptr;
  uint32_t* b_cnt = nullptr;
  uint32_t* s_base = nullptr;
  uint32_t* s_cnt = nullptr;
  uint32_t* s_depth = nullptr;  // max depth reached by the last Reset (statistics)
  uint32_t* b_depth = nullptr;
  float* b_sw = nullptr;
  float* b_sum = nullptr;
  float* b_area = nullptr;
  float* s_sw = nullptr;
  float* s_sum = nullptr;
  float* s_area = nullptr;
  float4* info = nullptr;
  float4* b_sums = nullptr;
  uint2* b_child = nullptr;
  float4* s_sums = nullptr;
  uint2* s_child = nullptr;
  uint4* s_node = nullptr;      // 2 per sampling node: sums bits, (children, 0, 0)
  uint4* s_jump = nullptr;      // pdf jump table: kPdfJumpCells per DTree
  uint2* b_jump = nullptr;      // record jump table of the building trees: kPdfJumpCells per DTree
};

// Kernel launchers (sdtree.hip).
hipError_t launch_dtree_build(DTreeArrays a, float* area_tmp, hipStream_t s);
// Per-DTree tracer view (info) and the interleaved copy of the sampling nodes.
hipError_t launch_dtree_publish(DTreeArrays a, hipStream_t s);
// Reset of the building trees of `dst` from the sampling trees of `src`
// (per new DTree j: source map[j], or j with map NULL); pass 0 counts nodes into dst.b_cnt,
// pass 1 writes children into dst.b_child at dst.b_base.
hipError_t launch_dtree_reset(DTreeArrays src, DTreeArrays dst, const uint32_t* map, uint32_t max_depth,
                              float subdiv_th, int pass, uint32_t* err, hipStream_t s);
// dst sampling trees := copies of src sampling trees (per map).
hipError_t launch_dtree_copy_sampling(DTreeArrays src, DTreeArrays dst, const uint32_t* map, hipStream_t s);
hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, hipStream_t s);

// Device refine of the spatial tree (RTSTreeT::Refine).  Counters in one
// array of 64-bit words that the host reads once per refine.
constexpr uint32_t kStatFootprint = 0;   // bytes of GetApproxMemoryFootPrint
constexpr uint32_t kStatLeaves = 1;      // leaves (DTrees) after the refine
constexpr uint32_t kStatBuildNodes = 2;  // building-tree nodes after the DTree reset
constexpr uint32_t kStatSampNodes = 3;   // sampling-tree nodes after the copies
constexpr uint32_t kStatErr = 4;         // error bits (low word)
constexpr uint32_t kStatWords = 5;
constexpr uint32_t kErrResetStack = 2u;  // DTree reset stack bound
constexpr uint32_t kErrStreeSize = 4u;   // a leaf would split more than 2^30 ways

// Per-node temporaries of the rebuild (n_new entries each; total: one word).
struct StreeScratch {
  uint32_t* is_leaf;
  uint32_t* rank;
  uint32_t* src;
  uint32_t* depth;
  uint32_t* total;
};

hipError_t launch_stree_footprint(const uint32_t* stree, uint32_t n, const uint32_t* b_cnt, const uint32_t* s_cnt,
                                  unsigned long long* stat, hipStream_t s);
// Split counts of the old leaves in walk order (order), their b^k in leaves.
hipError_t launch_stree_split_count(const uint32_t* stree, const uint32_t* order, uint32_t n_leaves,
                                    const float* b_sw, const uint32_t* reset_cnt, const uint32_t* s_cnt, float th,
                                    bool octree, int max_mb, unsigned long long* stat, uint32_t* leaves,
                                    uint32_t* splits, hipStream_t s);
// New node array (n_new words), walk order of its leaves, source DTree and leaf
// depth of every new DTree.
hipError_t launch_stree_rebuild(const uint32_t* stree_old, uint32_t n_old, const uint32_t* order_old,
                                uint32_t n_leaves_old, const uint32_t* leaves_excl, const uint32_t* splits,
                                const uint32_t* leaf_depth_old, bool octree, uint32_t n_new, uint32_t n_leaves,
                                StreeScratch t, uint32_t* stree_new, uint32_t* order_new, uint32_t* map,
                                uint32_t* leaf_depth_new, hipStream_t s);
hipError_t launch_dtree_gather_counts(const uint32_t* map, uint32_t n, const uint32_t* s_cnt_old,
                                      const uint32_t* reset_cnt, uint32_t* s_cnt, uint32_t* b_cnt, hipStream_t s);
hipError_t launch_stree_jump(const uint32_t* stree, uint32_t G, uint32_t* jump, hipStream_t s);

// Test/driver kernels on an SdTreeView.
hipError_t launch_sdtree_record(SdTreeView v, const float* pos, const float* canon, const float* irr,
                                const float* sw, uint32_t n, uint64_t seed, uint32_t batch, hipStream_t s);
hipError_t launch_sdtree_sample(SdTreeView v, const uint32_t* cells, uint32_t n, uint64_t seed, float* out_canon,
                                hipStream_t s);
hipError_t launch_sdtree_pdf(SdTreeView v, const uint32_t* cells, const float* canon, uint32_t n, float* out_pdf,
                             hipStream_t s);
hipError_t launch_sdtree_lookup(SdTreeView v, const float* pos, uint32_t n, uint32_t* out_cells, hipStream_t s);

}  // namespace rtpg
