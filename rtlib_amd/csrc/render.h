// Wavefront path state and the render launchers (host side of render.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_dev.h"
#include "views.h"

namespace rtpg {

constexpr float kRayTmin = 1e-4f;   // SimpleTrace.cu:76
// Traversal stack entries kept in LDS per lane; deeper entries spill to HBM.
constexpr uint32_t kShortStack = 8;
// The queues of live paths and shadow rays are split into
// regions of PathSoA::qreg slots (a multiple of 1024), each with its own
// counter per depth (count[depth * kQueueRegions + region]): a work-group
// appends to the region that holds its own slots.  Appends to one counter
// serialise (about 11 ns each on MI355X); with one counter per queue the
// one-atomic-per-work-group appends of a 32M-path bounce took 0.7-1.4 ms.
constexpr uint32_t kQueueRegions = 256;
constexpr float kRayTmax = 1e20f;

// Path meta word (ray_d.w): depth in bits 0..7, flags above.
constexpr uint32_t kMetaDelta = 0x100u;   // previous event was delta (or the camera)

// Guide vertex flags (vC.w).
constexpr uint32_t kVtxValid = 1u;
constexpr uint32_t kVtxDelta = 2u;
constexpr uint32_t kVtxBadBsdf = 4u;

// Structure-of-arrays path state for one launch chunk of n paths.
// Path pid = pixel + sample * npix.
struct PathSoA {
  uint32_t n;
  // The state of a live path moves with it: it is stored at the path's slot
  // in the queue of its depth (buffer depth & 1) and shading writes it to the
  // slot the path takes in the next queue, so every bounce kernel reads and
  // writes its rays contiguously however few paths are left.
  float4* ray_o[2];  // origin, w = pdf of the sampled direction (MIS)
  float4* ray_d[2];  // direction, w = meta word
  float4* T[2];      // throughput
  uint2* rng[2];     // PCG32 state
  float4* hit;       // per slot of the current depth: t, u, v, leaf-order triangle (bits; 0xFFFFFFFF = miss)
  uint32_t* queue[2];  // path index (pid) per slot
  uint32_t qreg;       // slots per queue region
  uint32_t nreg;       // regions in use: ceil(n / qreg)
  float4* L;         // radiance, per path (pid)
  uint32_t* nhit;    // hits processed, per path
  uint32_t* count;   // live paths per depth and queue region [(max_depth + 2) x kQueueRegions]
  uint32_t* scount;  // shadow rays per depth and queue region [(max_depth + 2) x kQueueRegions]
  float4* shadow;    // 3 x float4 per shadow ray
  float* moments;    // per pixel: sum and sum of squares of the samples' RGB mean, or NULL
  uint32_t* deep;    // traversal spill stack per ray slot: (entry, distance) pairs, 2 x (cap - kShortStack) x n words
  uint32_t* err;     // sticky error bits
  // guide vertices, [depth][pid]
  float4* vA;        // throughput after the bounce, woPdf
  float4* vB;        // contribution at this hit, cos of the sampled direction
  float4* vC;        // canonical direction (u, v), cell, flags
  // traversal work counters (null unless counting): extend nodes, extend
  // triangles, shadow nodes, shadow triangles
  unsigned long long* work;
  // path-replay adjoint (grad_alb non-null selects the adjoint kernels)
  uint32_t adj_mode;     // 0 primal pass, 1 replay and scatter
  uint32_t npix;
  float inv_spp;
  const float* dldi;     // dLoss/dimage, 3 per pixel
  const float4* adjL;    // primal radiance per path
  float* grad_alb;       // 3 per leaf-order triangle
  float* grad_le;        // 3 per light
  // guide vertices (vA, vB, shadow contributions) are recorded for pid <
  // rec_limit, or with the learned guide (mlp_stride > 0) for the paths
  // pid = slot * mlp_stride + mlp_phase
  uint32_t rec_limit;
  // learned guide: mixture per queue slot (kMlpMixF4 float4 + c), training
  // vertices of the recorded paths ([5][depth][slot] float4, mlp_rec slots)
  // and the sample buffer they are turned into
  float4* mix;
  float* mixc;
  float4* mv;
  uint32_t mlp_rec;
  uint32_t mlp_stride;
  uint32_t mlp_phase;
  uint32_t mlp_guide_depth;  // vertices at depth < this are guided and recorded
  // spatial cache of the learned guide (null: the per-slot mixtures of the
  // inference pass): entries of mlp_cache_entry over the network's box
  const float4* mcache;
  const float* mcache_c;
  uint32_t mcache_res;
  float mcache_lo[3];
  float mcache_inv;
  float4* mlp_out;
  uint32_t* mlp_count;
  uint32_t mlp_cap;
};

constexpr int kTimedKernels = 5;   // extend, shade, shadow, accumulate, commit
constexpr int kMaxDepthEv = 64;

hipError_t launch_trace_rays(const SceneView& sc, const float4* rays, uint32_t n, float4* hits, uint32_t* occl,
                             bool any, uint32_t* deep, uint32_t* err, hipStream_t s);

// Records this device's address of the allocation-free ray queries' scratch
// array (trace_noalloc.hip); called once per context creation, device current.
void scratch_setup(int device);

hipError_t launch_tonemap(const float* accum, uint32_t npix, float inv_spp, uchar4* frame, hipStream_t s);
// Per-work-group (256 pixels) sums of the relative variance of a moments render
// (n samples per pixel) against the accumulated image (acc_spp samples).
hipError_t launch_relvar(const float* mom, const float* accum, uint32_t npix, uint32_t n, uint32_t acc_spp,
                         double* partial, hipStream_t s);

hipError_t launch_debug(const RenderParams& rp, const Camera& cam, const SceneView& sc, const HashGridView& hg,
                        const MortonView& mq, uint32_t type, bool use_grid, float4* val, uchar4* frame,
                        uint32_t* deep, uint32_t* err, hipStream_t s);

hipError_t launch_render_chunk(const RenderParams& rp, const Camera& cam, const SceneView& sc, PathSoA ps,
                               const SdTreeView& sd, const HashGridView& hg, const MortonView& mq, const MlpView& ml,
                               int guide, float* accum, hipStream_t s, hipEvent_t* ev = nullptr);

}  // namespace rtpg
