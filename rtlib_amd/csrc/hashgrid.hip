// Hash grid + Morton direction quadtrees: device kernels.
#include <hip/hip_runtime.h>

#include "common.h"
#include "guide_dev.h"
#include "hashgrid.h"

namespace rtpg {

static inline uint32_t nblk(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

// Cell lookup for a batch of positions: read-only probe of the previous table
// (FindFromPrv), or insert-or-find in the current table (FindFromCur).
__global__ void k_grid_find(HashGridView h, const float* pos, uint32_t n, uint32_t insert, uint32_t* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t ix, iy, iz, b, c;
  grid_cell(h, V3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, ix, iy, iz);
  grid_hash(h, ix, iy, iz, b, c);
  out[i] = insert ? grid_find_cur(h, b, c) : grid_find_prv(h, b, c);
}

// mortonBuildKernel (OPX7UtilsMortonBuild.cu:2-16), one wave per cell: level
// by level from the leaves up, each parent adds its four children in code
// order, which is the order the reference's serial loop adds them in.
__global__ void __launch_bounds__(64) k_morton_build(float* w, uint32_t level, uint32_t per_cell, uint32_t cells) {
  uint32_t cell = blockIdx.x;
  if (cell >= cells) return;
  float* base = w + (size_t)per_cell * cell;
  for (uint32_t i = 0; i < level; ++i) {
    uint32_t src_lvl = level - i;
    uint32_t src = mq_level_offset(src_lvl);
    uint32_t dst = mq_level_offset(src_lvl - 1u);
    uint32_t parents = 1u << (2u * (src_lvl - 1u));
    for (uint32_t p = threadIdx.x; p < parents; p += 64u) {
      float v = base[dst + p];
      v += base[src + 4u * p + 0u];
      v += base[src + 4u * p + 1u];
      v += base[src + 4u * p + 2u];
      v += base[src + 4u * p + 3u];
      base[dst + p] = v;
    }
    __syncthreads();
  }
}

// The same for level <= 3 (at most 16 parents per step), 16 lanes per cell:
// lane p adds parent p of the deepest step from its four leaves (the 16 lanes
// read 64 consecutive floats), and each coarser step takes the four children
// from the lanes that hold them; the additions keep the order above (parent,
// then children 0..3).  Four cells per wave instead of one.
__global__ void __launch_bounds__(256) k_morton_build16(float* w, uint32_t level, uint32_t per_cell,
                                                         uint32_t cells) {
  const uint32_t lane = threadIdx.x & 15u;
  const uint32_t cell = (blockIdx.x * 256u + threadIdx.x) >> 4;
  const bool in = cell < cells;  // uniform over the 16 lanes of a cell
  float* base = w + (size_t)per_cell * (in ? cell : 0u);
  float cur = 0.0f;
  for (uint32_t src_lvl = level; src_lvl >= 1u; --src_lvl) {
    const uint32_t src = mq_level_offset(src_lvl), dst = mq_level_offset(src_lvl - 1u);
    const uint32_t parents = 1u << (2u * (src_lvl - 1u));
    float c[4];
    if (src_lvl == level) {
#pragma unroll
      for (int k = 0; k < 4; ++k) c[k] = (in && lane < parents) ? base[src + 4u * lane + (uint32_t)k] : 0.0f;
    } else {
      // lanes 4p .. 4p + 3 hold the children of parent p (4p + 3 <= 15)
#pragma unroll
      for (int k = 0; k < 4; ++k) c[k] = __shfl(cur, (int)((4u * lane + (uint32_t)k) & 15u), 16);
    }
    if (in && lane < parents) {
      float v = base[dst + lane];
      v += c[0];
      v += c[1];
      v += c[2];
      v += c[3];
      base[dst + lane] = v;
      cur = v;
    }
  }
}

// MortonQuadTreeT::Record leaf-only branch (OPX7UtilsMorton.h:262-285) for a batch.
__global__ void k_morton_record(MortonView m, const uint32_t* cells, const float* canon, const float* val,
                                uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t cell = cells[i];
  if (cell == kInvalidCell) return;
  uint32_t nn = 1u << m.level;
  uint32_t x = (uint32_t)(canon[2 * i] * (float)nn), y = (uint32_t)(canon[2 * i + 1] * (float)nn);
  x = x < nn ? x : nn - 1u;
  y = y < nn ? y : nn - 1u;
  unsafeAtomicAdd(m.build + (size_t)m.per_cell * cell + mq_level_offset(m.level) + morton2(x, y), val[i]);
}

__global__ void k_morton_sample(MortonView m, const uint32_t* cells, uint32_t n, uint64_t seed, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pcg32 rng = pcg32_seed(seed, i);
  float u, v, pdf;
  // chain b evaluates a point spread over the square beside the sampling
  // chain; a lone evaluation must give the same bits, else the pdf is NaN
  float bx = check_coord(i, 2654435769u), by = check_coord(i, 0x85EBCA6Bu), pb, pb1, unused;
  mq_pair(m, cells[i], true, rng, u, v, pdf, true, bx, by, pb);
  mq_pair(m, cells[i], false, rng, bx, by, pb1, false, 0.0f, 0.0f, unused);
  if (__float_as_uint(pb) != __float_as_uint(pb1)) pdf = __uint_as_float(0x7FC00000u);
  out[3 * i] = u;
  out[3 * i + 1] = v;
  out[3 * i + 2] = pdf;
}

__global__ void k_morton_pdf(MortonView m, const uint32_t* cells, const float* canon, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pcg32 rng = pcg32_seed(0u, i);  // drawn from by neither chain
  float u = canon[2 * i], v = canon[2 * i + 1], pa, pb;
  mq_pair(m, cells[i], false, rng, u, v, pa, true, canon[2 * i], canon[2 * i + 1], pb);
  out[i] = __float_as_uint(pa) == __float_as_uint(pb) ? pa : __uint_as_float(0x7FC00000u);
}

__global__ void k_fill_zero(float* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0f;
}

hipError_t launch_grid_find(HashGridView h, const float* pos, uint32_t n, bool insert, uint32_t* out,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grid_find, dim3(nblk(n, 256)), dim3(256), 0, s, h, pos, n, insert ? 1u : 0u, out);
  return hipGetLastError();
}
hipError_t launch_morton_build(float* w, uint32_t level, uint32_t per_cell, uint32_t cells, hipStream_t s) {
  if (cells == 0 || level == 0) return hipSuccess;
  if (level <= 3u) {
    hipLaunchKernelGGL(k_morton_build16, dim3(nblk(cells, 16u)), dim3(256), 0, s, w, level, per_cell, cells);
  } else {
    hipLaunchKernelGGL(k_morton_build, dim3(cells), dim3(64), 0, s, w, level, per_cell, cells);
  }
  return hipGetLastError();
}
hipError_t launch_morton_record(MortonView m, const uint32_t* cells, const float* canon, const float* val,
                                uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_record, dim3(nblk(n, 256)), dim3(256), 0, s, m, cells, canon, val, n);
  return hipGetLastError();
}
hipError_t launch_morton_sample(MortonView m, const uint32_t* cells, uint32_t n, uint64_t seed, float* out,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_sample, dim3(nblk(n, 256)), dim3(256), 0, s, m, cells, n, seed, out);
  return hipGetLastError();
}
hipError_t launch_morton_pdf(MortonView m, const uint32_t* cells, const float* canon, uint32_t n, float* out,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_pdf, dim3(nblk(n, 256)), dim3(256), 0, s, m, cells, canon, n, out);
  return hipGetLastError();
}
hipError_t launch_fill_zero(float* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill_zero, dim3((uint32_t)blocks), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace rtpg
