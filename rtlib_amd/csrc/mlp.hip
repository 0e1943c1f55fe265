// Learned directional guide: training step (forward + backward on the matrix
// cores, hash-grid gradient scatter), deterministic split-K gradient reduce,
// Adam, initialisation and the query kernels used by the C ABI.
#include "mlp.h"
#include "mlp_tile.h"
#include "sdtree.h"

namespace rtpg {

constexpr int kSP = 68;  // staging row pitch (bf16): 64 samples + 4

// Stages this wave's 16 samples of the layer input a (NBA feature blocks) and
// of the output gradient dz (NBD blocks) as [feature][sample] rows.
template <int NBA, int NBD>
__device__ __forceinline__ void stage_pair(uint16_t* sa, uint16_t* sd, const bf16x4_t* a, const bf16x4_t* dz,
                                           int wave, int lane) {
  const int r = lane & 15, g = lane >> 4, col = 16 * wave + r;
#pragma unroll
  for (int b = 0; b < NBA; ++b)
#pragma unroll
    for (int j = 0; j < 4; ++j) sa[(16 * b + 4 * g + j) * kSP + col] = (uint16_t)a[b][j];
#pragma unroll
  for (int b = 0; b < NBD; ++b)
#pragma unroll
    for (int j = 0; j < 4; ++j) sd[(16 * b + 4 * g + j) * kSP + col] = (uint16_t)dz[b][j];
}

// dW = dz . a^T over the work-group's 64 staged samples; of the NT x NU
// output tiles this wave owns tiles wave, wave + 4, ... (NOWN of them).
template <int NT, int NU, int NOWN>
__device__ __forceinline__ void dw_tiles(const uint16_t* sa, const uint16_t* sd, f32x4_t* acc, int wave, int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < NOWN; ++i) {
    int k = wave + 4 * i;
    int t = k / NU, u = k % NU;
#pragma unroll
    for (int ks = 0; ks < 4; ks += 2) {
      const uint16_t* pa = sd + (16 * t + r) * kSP + 16 * ks + 4 * g;
      const uint16_t* pb = sa + (16 * u + r) * kSP + 16 * ks + 4 * g;
      acc[i] = mfma32(*reinterpret_cast<const bf16x4_t*>(pa), *reinterpret_cast<const bf16x4_t*>(pa + 16),
                      *reinterpret_cast<const bf16x4_t*>(pb), *reinterpret_cast<const bf16x4_t*>(pb + 16), acc[i]);
    }
  }
}

template <int NT, int NU, int NOWN, int IN>
__device__ __forceinline__ void dw_store(float* dst, const f32x4_t* acc, int wave, int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < NOWN; ++i) {
    int k = wave + 4 * i;
    int t = k / NU, u = k % NU;
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[(16 * t + 4 * g + j) * IN + 16 * u + r] = acc[i][j];
  }
}

// Spatial order of a training window: samples bucketed by the Morton code of
// their position on a 64^3 grid over the guide's box, so that the 16 samples
// of a wave row mostly share the cells of the coarse levels and the per-row
// run sums replace most of their table-gradient atomics.  The order inside a
// bucket is the atomic arrival order (a step's gradient is a sum).
__device__ __forceinline__ uint32_t mlp_sort_key(const MlpView& m, float4 s0) {
  const float n = (float)kMlpSortCells;
  float q[3] = {(s0.x - m.lo[0]) * m.inv_size * n, (s0.y - m.lo[1]) * m.inv_size * n,
                (s0.z - m.lo[2]) * m.inv_size * n};
  uint32_t c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = (uint32_t)fminf(fmaxf(floorf(q[k]), 0.0f), n - 1.0f);
  return morton3_30(c[0], c[1], c[2]);
}

__global__ void k_mlp_bucket_count(MlpView m, const float4* smp, uint32_t n, uint32_t* hist) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(hist + mlp_sort_key(m, smp[4u * i]), 1u);
}

// offs: exclusive prefix sums of the bucket counts (advanced by the scatter).
__global__ void k_mlp_bucket_scatter(MlpView m, const float4* smp, uint32_t n, uint32_t* offs, float4* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 s0 = smp[4u * i];
  uint32_t d = atomicAdd(offs + mlp_sort_key(m, s0), 1u);
  out[4u * d] = s0;
  out[4u * d + 1u] = smp[4u * i + 1u];
  out[4u * d + 2u] = smp[4u * i + 2u];
  out[4u * d + 3u] = smp[4u * i + 3u];
}

// One training pass over n samples: loss = mean_i -w_i log(c q(w_i) + (1 - c) p_bsdf_i)
// (the KL divergence of the mixture from the radiance-weighted target up to
// a constant).  Each work-group writes its weight-gradient sum (unscaled) and
// loss sum; the hash-table gradient is scattered with float atomics.
__global__ void __launch_bounds__(kMlpTrainBlock) k_mlp_train(MlpView m, const float4* smp, uint32_t n,
                                                              float* partial, float* loss_partial) {
  __shared__ uint16_t lw[kLdsWeights];
  __shared__ uint16_t sa[kMlpHidden * kSP];
  __shared__ uint16_t sd[kMlpHidden * kSP];
  __shared__ float hs[4][16 * kHeadPitch];
  __shared__ float lred[4];
  __shared__ uint32_t gkeys[kMlpGradTab];
  __shared__ float gvals[2 * kMlpGradTab];
  for (uint32_t i = threadIdx.x; i < kMlpGradTab; i += blockDim.x) {
    gkeys[i] = kMlpEmptyKey;
    gvals[2 * i] = 0.0f;
    gvals[2 * i + 1] = 0.0f;
  }
  const MlpGradTable tab{m.grad ? gkeys : nullptr, gvals};
  mlp_load_weights(lw, m.w);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  f32x4_t a0[3], a1[4], a2[4], a3[3];
  const f32x4_t z4 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    a1[i] = z4;
    a2[i] = z4;
    if (i < 3) {
      a0[i] = z4;
      a3[i] = z4;
    }
  }
  float lsum = 0.0f;
  float* row = hs[wave] + r * kHeadPitch;
  // a contiguous slice of the (spatially ordered) samples per work-group
  const uint32_t chunks = (n + 63u) / 64u;
  const uint32_t c_end = (uint32_t)(((uint64_t)chunks * (blockIdx.x + 1u)) / gridDim.x);
  for (uint32_t chunk = (uint32_t)(((uint64_t)chunks * blockIdx.x) / gridDim.x); chunk < c_end; ++chunk) {
    uint32_t idx = chunk * 64u + 16u * (uint32_t)wave + (uint32_t)r;
    bool valid = idx < n;
    MlpInput in{V3{0, 0, 0}, V3{0, 0, 1}, V3{0, 0, 1}, 0.0f};
    V3 om{0.0f, 0.0f, 1.0f};
    float w = 0.0f, pb = 1.0f;
    if (valid) {
      float4 s0 = smp[4u * idx], s1 = smp[4u * idx + 1u], s2 = smp[4u * idx + 2u], s3 = smp[4u * idx + 3u];
      in.p = V3{s0.x, s0.y, s0.z};
      in.rough = s0.w;
      in.n = V3{s1.x, s1.y, s1.z};
      pb = s1.w;
      in.wo = V3{s2.x, s2.y, s2.z};
      w = s2.w;
      om = V3{s3.x, s3.y, s3.z};
    }
    MlpActs act;
    mlp_inputs(m, in, lane, act.x0);
    f32x4_t o[3];
    mlp_forward(lw, act, o, lane);
    mlp_spill(hs[wave], o, lane);
    __syncthreads();
    MlpHead h = mlp_head(row, lane);
    float u[2], d[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      d[q] = dot(h.mu[q], om);
      u[q] = h.A[q] * expf(h.kappa[q] * (d[q] - 1.0f));
    }
    float qv = u[0] + u[1];
    qv += __shfl_xor(qv, 16);
    qv += __shfl_xor(qv, 32);
    float qp = fmaxf(h.c * qv + (1.0f - h.c) * pb, 1e-20f);
    float G = -w / qp;
    if (g == 0 && valid) lsum += -w * logf(qp);
    float dq = G * h.c;
    float dt = G * (qv - pb) * kMlpSelSpan * h.sig * (1.0f - h.sig);
    __syncthreads();  // every lane has read its raw outputs
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      float* oj = row + 5 * (2 * g + q);
      float k = h.kappa[q];
      oj[0] = dq * (u[q] - h.pi[q] * qv);
      float sm = dq * u[q] * k / h.mnorm[q];
      oj[1] = sm * (om.x - h.mu[q].x * d[q]);
      oj[2] = sm * (om.y - h.mu[q].y * d[q]);
      oj[3] = sm * (om.z - h.mu[q].z * d[q]);
      bool inside = h.sk[q] > kMlpLogKappaMin && h.sk[q] < kMlpLogKappaMax;
      float e = expf(-2.0f * k);
      float coth = -(1.0f + e) / expm1f(-2.0f * k);
      oj[4] = inside ? dq * u[q] * k * (1.0f / k - coth + d[q]) : 0.0f;
    }
    if (g == 0) {
      row[40] = dt;
      for (int k = 41; k < kMlpOut; ++k) row[k] = 0.0f;
    }
    __syncthreads();
    bf16x4_t dz[4];
#pragma unroll
    for (int t = 0; t < 3; ++t)
      dz[t] = pack_bf16(row[16 * t + 4 * g], row[16 * t + 4 * g + 1], row[16 * t + 4 * g + 2], row[16 * t + 4 * g + 3]);
    f32x4_t dh[4];
    // layer 3 (48 x 64)
    stage_pair<4, 3>(sa, sd, act.h3, dz, wave, lane);
    __syncthreads();
    dw_tiles<3, 4, 3>(sa, sd, a3, wave, lane);
    __syncthreads();
    dense_bwd<4, 3>(lw + kLdsW3, kLdsP1, dz, dh, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) dz[t] = relu_grad(dh[t], act.h3[t]);
    // layer 2 (64 x 64)
    stage_pair<4, 4>(sa, sd, act.h2, dz, wave, lane);
    __syncthreads();
    dw_tiles<4, 4, 4>(sa, sd, a2, wave, lane);
    __syncthreads();
    dense_bwd<4, 4>(lw + kLdsW2, kLdsP1, dz, dh, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) dz[t] = relu_grad(dh[t], act.h2[t]);
    // layer 1 (64 x 64)
    stage_pair<4, 4>(sa, sd, act.h1, dz, wave, lane);
    __syncthreads();
    dw_tiles<4, 4, 4>(sa, sd, a1, wave, lane);
    __syncthreads();
    dense_bwd<4, 4>(lw + kLdsW1, kLdsP1, dz, dh, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) dz[t] = relu_grad(dh[t], act.h1[t]);
    // layer 0 (64 x 48)
    stage_pair<3, 4>(sa, sd, act.x0, dz, wave, lane);
    __syncthreads();
    dw_tiles<4, 3, 3>(sa, sd, a0, wave, lane);
    __syncthreads();
    // gradient of the 32 encoded features: lane holds levels {2g, 2g+1} and {8+2g, 9+2g}
    dense_bwd<2, 4>(lw + kLdsW0, kLdsP0, dz, dh, lane);
    {
      float px = fminf(fmaxf((in.p.x - m.lo[0]) * m.inv_size, 0.0f), 1.0f);
      float py = fminf(fmaxf((in.p.y - m.lo[1]) * m.inv_size, 0.0f), 1.0f);
      float pz = fminf(fmaxf((in.p.z - m.lo[2]) * m.inv_size, 0.0f), 1.0f);
      // levels below 1024 cells per axis: neighbouring samples (the window
      // is in spatial order) share cells, so each row of 16 samples sums its
      // runs before the atomics (fp32 through the work-group table for dense
      // levels and those below kMlpPkFirst, packed bf16 pairs for the other
      // hashed levels; the choice is uniform over a row: one level per row);
      // finer levels send every sample's contributions directly
#pragma unroll
      for (int b = 0; b < 2; ++b) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int l = 8 * b + 2 * g + q;
          const float g0 = dh[b][2 * q], g1 = dh[b][2 * q + 1];
          if (m.res[l] < 1024.0f) {
            if (l >= kMlpPkFirst && !((m.dense >> l) & 1u)) {
              mlp_encode_level_grad_runs<true>(m, l, px, py, pz, g0, g1, valid);
            } else {
              mlp_encode_level_grad_runs<false>(m, l, px, py, pz, g0, g1, valid, tab);
            }
          } else if (valid) {
            if (l >= kMlpPkFirst) {
              mlp_encode_level_grad_pk(m, l, px, py, pz, g0, g1);
            } else {
              mlp_encode_level_grad(m, l, px, py, pz, g0, g1);
            }
          }
        }
      }
    }
  }
  if (m.grad) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kMlpGradTab; i += blockDim.x) {
      const uint32_t key = gkeys[i];
      if (key == kMlpEmptyKey) continue;
      float* gt = m.grad + (size_t)key * 2u;
      if (gvals[2 * i] != 0.0f) unsafeAtomicAdd(gt, gvals[2 * i]);
      if (gvals[2 * i + 1] != 0.0f) unsafeAtomicAdd(gt + 1, gvals[2 * i + 1]);
    }
  }
  float* dst = partial + (size_t)blockIdx.x * kMlpWeights;
  dw_store<4, 3, 3, kMlpIn>(dst + kMlpW0, a0, wave, lane);
  dw_store<4, 4, 4, kMlpHidden>(dst + kMlpW1, a1, wave, lane);
  dw_store<4, 4, 4, kMlpHidden>(dst + kMlpW2, a2, wave, lane);
  dw_store<3, 4, 3, kMlpHidden>(dst + kMlpW3, a3, wave, lane);
  for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0) lred[wave] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = lred[0] + lred[1] + lred[2] + lred[3];
}

// Sums the per-work-group partials in a fixed order (deterministic): a
// work-group of 4 waves owns 64 weights (one per lane); wave w adds the
// partial rows w, w + 4, w + 8, ... and the four wave sums are added in wave
// order.  Work-group 0 also sums the loss partials the same way (256 lanes,
// then a fixed tree).
constexpr uint32_t kReduceCols = 64;

__global__ void __launch_bounds__(256) k_mlp_reduce(const float* partial, const float* loss_partial, uint32_t grid,
                                                    float* grad_w, float* loss) {
  __shared__ float red[4][kReduceCols];
  __shared__ float lred[256];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t k = blockIdx.x * kReduceCols + lane;
  float s = 0.0f;
  if (k < (uint32_t)kMlpWeights) {
    for (uint32_t b = wave; b < grid; b += 4u) s += partial[(size_t)b * kMlpWeights + k];
  }
  red[wave][lane] = s;
  if (blockIdx.x == 0) {
    float t = 0.0f;
    for (uint32_t b = threadIdx.x; b < grid; b += 256u) t += loss_partial[b];
    lred[threadIdx.x] = t;
  }
  __syncthreads();
  if (wave == 0 && k < (uint32_t)kMlpWeights) grad_w[k] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
  if (blockIdx.x == 0) {
    for (uint32_t h = 128; h >= 1; h >>= 1) {
      if (threadIdx.x < h) lred[threadIdx.x] += lred[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) *loss = lred[0];
  }
}

__device__ __forceinline__ float adam_one(float& p, float& m1, float& m2, float gg, float lr_t, float b1, float b2,
                                          float eps) {
  m1 = b1 * m1 + (1.0f - b1) * gg;
  m2 = b2 * m2 + (1.0f - b2) * gg * gg;
  p -= lr_t * m1 / (sqrtf(m2) + eps);
  return p;
}

__global__ void k_mlp_adam(float* p, float* m1, float* m2, float* g, float gscale, size_t n, float lr_t, float b1,
                           float b2, float eps, uint16_t* w_bf16, uint32_t* table_bf16, uint32_t* gpk,
                           size_t pk_first) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (table_bf16) {
    size_t e = 2 * i;
    if (e >= n) return;
    // table entry i: fp32 gradient, plus the packed bf16 pair from pk_first on
    float gp[2] = {0.0f, 0.0f};
    if (gpk && i >= pk_first) {
      uint32_t u = gpk[i - pk_first];
      gpk[i - pk_first] = 0u;
      gp[0] = bf2f((uint16_t)(u & 0xFFFFu));
      gp[1] = bf2f((uint16_t)(u >> 16));
    }
    float v[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      float gg = (g[e + k] + gp[k]) * gscale;
      g[e + k] = 0.0f;
      float pp = p[e + k], a = m1[e + k], b = m2[e + k];
      v[k] = adam_one(pp, a, b, gg, lr_t, b1, b2, eps);
      p[e + k] = pp;
      m1[e + k] = a;
      m2[e + k] = b;
    }
    table_bf16[i] = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
  } else {
    if (i >= n) return;
    float gg = g[i] * gscale;
    g[i] = 0.0f;
    float pp = p[i], a = m1[i], b = m2[i];
    adam_one(pp, a, b, gg, lr_t, b1, b2, eps);
    p[i] = pp;
    m1[i] = a;
    m2[i] = b;
    w_bf16[i] = f2bf(pp);
  }
}

// Weights: uniform(+-sqrt(6 / fan_in)) (output layer x 0.1); the zero input
// columns of W0 and the unused output rows of W3 start (and stay) at zero.
__global__ void k_mlp_init_w(float* wm, uint16_t* wb, uint64_t seed) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kMlpWeights) return;
  int layer, row, col, in;
  if (k < kMlpW1) {
    layer = 0;
    in = kMlpIn;
    row = k / in;
    col = k % in;
  } else {
    int e = k < kMlpW2 ? k - kMlpW1 : (k < kMlpW3 ? k - kMlpW2 : k - kMlpW3);
    layer = k < kMlpW2 ? 1 : (k < kMlpW3 ? 2 : 3);
    in = kMlpHidden;
    row = e / in;
    col = e % in;
  }
  Pcg32 rng = pcg32_seed(seed, (uint64_t)k);
  float fan = layer == 0 ? 40.0f : (float)kMlpHidden;
  float bound = sqrtf(6.0f / fan) * (layer == 3 ? 0.1f : 1.0f);
  float v = (2.0f * rnd(rng) - 1.0f) * bound;
  if ((layer == 0 && col >= 40) || (layer == 3 && row >= kMlpOutUsed)) v = 0.0f;
  wm[k] = v;
  wb[k] = f2bf(v);
}

__global__ void k_mlp_init_table(float* t, uint32_t* tb, size_t pairs, uint64_t seed) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pairs) return;
  Pcg32 rng = pcg32_seed(seed ^ 0x5bd1e995ULL, (uint64_t)i);
  float a = (2.0f * rnd(rng) - 1.0f) * 1e-4f;
  float b = (2.0f * rnd(rng) - 1.0f) * 1e-4f;
  t[2 * i] = a;
  t[2 * i + 1] = b;
  tb[i] = (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}

// Queries: raw[41] outputs and the mixture (pi, mu, kappa per lobe, then c).
__global__ void __launch_bounds__(256) k_mlp_forward_q(MlpView m, const float4* q, uint32_t n, float* raw,
                                                       float* mixo) {
  __shared__ uint16_t lw[kLdsWeights];
  __shared__ float hs[4][16 * kHeadPitch];
  mlp_load_weights(lw, m.w);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  uint32_t idx = blockIdx.x * 64u + 16u * (uint32_t)wave + (uint32_t)r;
  bool valid = idx < n;
  MlpInput in{V3{0, 0, 0}, V3{0, 0, 1}, V3{0, 0, 1}, 0.0f};
  if (valid) {
    float4 a = q[3u * idx], b = q[3u * idx + 1u], c = q[3u * idx + 2u];
    in.p = V3{a.x, a.y, a.z};
    in.rough = a.w;
    in.n = V3{b.x, b.y, b.z};
    in.wo = V3{c.x, c.y, c.z};
  }
  MlpActs act;
  mlp_inputs(m, in, lane, act.x0);
  f32x4_t o[3];
  mlp_forward(lw, act, o, lane);
  mlp_spill(hs[wave], o, lane);
  __syncthreads();
  const float* row = hs[wave] + r * kHeadPitch;
  MlpHead h = mlp_head(row, lane);
  if (!valid) return;
  if (raw) {
    for (int k = g; k < kMlpOutUsed; k += 4) raw[(size_t)idx * kMlpOutUsed + k] = row[k];
  }
  if (mixo) {
    float* mo = mixo + (size_t)idx * kMlpOutUsed;
#pragma unroll
    for (int qq = 0; qq < 2; ++qq) {
      int j = 2 * g + qq;
      mo[5 * j + 0] = h.pi[qq];
      mo[5 * j + 1] = h.mu[qq].x;
      mo[5 * j + 2] = h.mu[qq].y;
      mo[5 * j + 3] = h.mu[qq].z;
      mo[5 * j + 4] = h.kappa[qq];
    }
    if (g == 0) mo[40] = h.c;
  }
}

// Spatial cache fill: every entry's mixture (kMlpMixF4 float4, the layout
// k_shade reads) and selection probability, 64 entries per work-group.
__global__ void __launch_bounds__(256) k_mlp_cache_fill(MlpView m, uint32_t n, float4* mix, float* cout) {
  __shared__ uint16_t lw[kLdsWeights];
  __shared__ float hs[4][16 * kHeadPitch];
  mlp_load_weights(lw, m.w);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15;
  uint32_t e = blockIdx.x * 64u + 16u * (uint32_t)wave + (uint32_t)r;
  bool valid = e < n;
  MlpInput in = mlp_cache_input(m, valid ? e : 0u);
  MlpActs act;
  mlp_inputs(m, in, lane, act.x0);
  f32x4_t o[3];
  mlp_forward(lw, act, o, lane);
  mlp_spill(hs[wave], o, lane);
  __syncthreads();
  MlpHead h = mlp_head(hs[wave] + r * kHeadPitch, lane);
  if (valid) mlp_write_mix(h, reinterpret_cast<float*>(mix + (size_t)e * kMlpMixF4), cout + e, lane);
}

__global__ void k_mlp_sample_q(const float* mix, uint32_t n, uint64_t seed, float4* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* mo = mix + (size_t)i * kMlpOutUsed;
  float4 mx[kMlpMixF4];
  float* f = reinterpret_cast<float*>(mx);
  for (int j = 0; j < kMlpLobes; ++j) {
    f[4 * j + 0] = mo[5 * j + 1];
    f[4 * j + 1] = mo[5 * j + 2];
    f[4 * j + 2] = mo[5 * j + 3];
    f[4 * j + 3] = mo[5 * j + 4];
    f[32 + j] = mo[5 * j];
    f[40 + j] = mo[5 * j] * vmf_norm(mo[5 * j + 4]);
  }
  Pcg32 rng = pcg32_seed(seed, (uint64_t)i);
  V3 d = mix_sample(mx, rng);
  out[i] = make_float4(d.x, d.y, d.z, mix_pdf(mx, d));
}

// ------------------------------------------------------------------ launchers
static inline uint32_t nblk(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

hipError_t launch_mlp_init(float* w_master, uint16_t* w_bf16, float* table, uint32_t* table_bf16, size_t table_floats,
                           uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_init_w, dim3(nblk(kMlpWeights, 256)), dim3(256), 0, s, w_master, w_bf16, seed);
  size_t pairs = table_floats / 2;
  hipLaunchKernelGGL(k_mlp_init_table, dim3(nblk(pairs, 256)), dim3(256), 0, s, table, table_bf16, pairs, seed);
  return hipGetLastError();
}

hipError_t launch_mlp_train(const MlpView& m, const float4* samples, uint32_t n, float* partial, float* loss_partial,
                            uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(kMlpTrainBlock), 0, s, m, samples, n, partial, loss_partial);
  return hipGetLastError();
}

hipError_t launch_mlp_spatial_order(const MlpView& m, const float4* samples, uint32_t n, uint32_t* hist,
                                    uint32_t* offs, uint32_t* total, float4* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t buckets = kMlpSortCells * kMlpSortCells * kMlpSortCells;
  hipError_t e = hipMemsetAsync(hist, 0, buckets * sizeof(uint32_t), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_bucket_count, dim3(nblk(n, 256)), dim3(256), 0, s, m, samples, n, hist);
  e = launch_exclusive_scan(hist, offs, buckets, total, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_bucket_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, m, samples, n, offs, out);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce(const float* partial, const float* loss_partial, uint32_t grid, float* grad_w,
                             float* loss, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3(nblk(kMlpWeights, kReduceCols)), dim3(256), 0, s, partial, loss_partial, grid,
                     grad_w, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* p, float* m1, float* m2, float* g, float gscale, size_t n, float lr_t, float b1,
                           float b2, float eps, uint16_t* w_bf16, uint32_t* table_bf16, uint32_t* gpk,
                           size_t pk_first, hipStream_t s) {
  size_t threads = table_bf16 ? n / 2 : n;
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(threads, 256)), dim3(256), 0, s, p, m1, m2, g, gscale, n, lr_t, b1, b2,
                     eps, w_bf16, table_bf16, gpk, pk_first);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const MlpView& m, const float4* q, uint32_t n, float* raw, float* mix, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward_q, dim3(nblk(n, 64)), dim3(256), 0, s, m, q, n, raw, mix);
  return hipGetLastError();
}

hipError_t launch_mlp_cache_fill(const MlpView& m, uint32_t n, float4* mix, float* c, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_cache_fill, dim3(nblk(n, 64)), dim3(256), 0, s, m, n, mix, c);
  return hipGetLastError();
}

hipError_t launch_mlp_sample(const float* mix, uint32_t n, uint64_t seed, float4* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_sample_q, dim3(nblk(n, 256)), dim3(256), 0, s, mix, n, seed, out);
  return hipGetLastError();
}

}  // namespace rtpg
