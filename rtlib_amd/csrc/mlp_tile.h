// Wave-level MLP tiles on the gfx950 matrix cores: v_mfma_f32_16x16x32_bf16
// for pairs of 16-wide k blocks, v_mfma_f32_16x16x16_bf16 for a lone block.
//
// Operand layout of the 16-wide form, lane l = 16 g + r (r < 16, g < 4):
//   A (16 x 16, rows x k): lane holds A[r][4g + j], j < 4
//   B (16 x 16, k x cols): lane holds B[4g + j][r]
//   D (16 x 16):           lane holds D[4g + j][r]
// The 32-wide form takes 8 k values per lane on A and B.  A lane gives it
// its 4 values of block 2b followed by its 4 of block 2b + 1, on A and B
// alike: the sum over k is the same whichever k a lane holds as long as its A
// and B values pair up, so two blocks fuse into one instruction without
// moving data.  With weights as A and activations as B (columns = 16
// samples), the D tile of one layer holds, per lane, sample r and features
// 4g + j of a 16-feature block: the B operand of the next layer.  Device code
// only.
#pragma once
#include "mlp_dev.h"

namespace rtpg {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef short bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_v __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// LDS copy of the weights: row pitches padded by 8 bf16 against bank conflicts.
constexpr int kLdsP0 = kMlpIn + 8;       // W0: 64 rows
constexpr int kLdsP1 = kMlpHidden + 8;   // W1, W2: 64 rows; W3: 48 rows
constexpr int kLdsW0 = 0;
constexpr int kLdsW1 = kLdsW0 + kMlpHidden * kLdsP0;
constexpr int kLdsW2 = kLdsW1 + kMlpHidden * kLdsP1;
constexpr int kLdsW3 = kLdsW2 + kMlpHidden * kLdsP1;
constexpr int kLdsWeights = kLdsW3 + kMlpOut * kLdsP1;  // 16256 bf16
constexpr int kHeadPitch = 49;                           // floats per sample row of the head scratch

// 19 wait states, with nothing scheduled across them.
__device__ __forceinline__ void mfma_drain() {
  __builtin_amdgcn_sched_barrier(0);
  asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 3");
  __builtin_amdgcn_sched_barrier(0);
}

// The lone 16-wide block of a layer accumulates onto the 32-wide form's
// result.  Issued right behind that v_mfma_f32_16x16x32_bf16 (as the compiler
// schedules it, with no wait between the two), the 16x16x16 form read a
// stale accumulator now and then: the network's outputs for identical inputs
// and weights differed from run to run in whole waves.  The wait before it
// makes them identical.
__device__ __forceinline__ f32x4_t mfma16(bf16x4_t a, bf16x4_t b, f32x4_t c) {
  mfma_drain();
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two 16-wide k blocks in one instruction: A = (a0 | a1), B = (b0 | b1).
__device__ __forceinline__ f32x4_t mfma32(bf16x4_t a0, bf16x4_t a1, bf16x4_t b0, bf16x4_t b1, f32x4_t c) {
  const bf16x8_t a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
  const bf16x8_t b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_v, a), __builtin_bit_cast(bf16x8_v, b), c,
                                                 0, 0, 0);
}

// Copies the row-major bf16 weights into the padded LDS layout (whole block).
__device__ __forceinline__ void mlp_load_weights(uint16_t* lw, const uint16_t* gw) {
  const uint32_t* g32 = reinterpret_cast<const uint32_t*>(gw);
  for (int p = threadIdx.x; p < kMlpWeights / 2; p += blockDim.x) {
    int e = 2 * p;
    int dst;
    if (e < kMlpW1) {
      dst = kLdsW0 + (e / kMlpIn) * kLdsP0 + e % kMlpIn;
    } else if (e < kMlpW2) {
      e -= kMlpW1;
      dst = kLdsW1 + (e / kMlpHidden) * kLdsP1 + e % kMlpHidden;
    } else if (e < kMlpW3) {
      e -= kMlpW2;
      dst = kLdsW2 + (e / kMlpHidden) * kLdsP1 + e % kMlpHidden;
    } else {
      e -= kMlpW3;
      dst = kLdsW3 + (e / kMlpHidden) * kLdsP1 + e % kMlpHidden;
    }
    *reinterpret_cast<uint32_t*>(lw + dst) = g32[p];
  }
}

__device__ __forceinline__ bf16x4_t pack_bf16(float a, float b, float c, float d) {
  bf16x4_t v;
  v[0] = (short)f2bf(a);
  v[1] = (short)f2bf(b);
  v[2] = (short)f2bf(c);
  v[3] = (short)f2bf(d);
  return v;
}

__device__ __forceinline__ bf16x4_t relu_bf16(f32x4_t y) {
  return pack_bf16(fmaxf(y[0], 0.0f), fmaxf(y[1], 0.0f), fmaxf(y[2], 0.0f), fmaxf(y[3], 0.0f));
}

// y[t] = W[16t .. 16t+15][:] . x for NT output tiles over NKB 16-wide k blocks.
template <int NT, int NKB>
__device__ __forceinline__ void dense_fwd(const uint16_t* W, int pitch, const bf16x4_t* x, f32x4_t* y, int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* row = W + (16 * t + r) * pitch + 4 * g;
#pragma unroll
    for (int kb = 0; kb + 1 < NKB; kb += 2) {
      bf16x4_t a0 = *reinterpret_cast<const bf16x4_t*>(row + 16 * kb);
      bf16x4_t a1 = *reinterpret_cast<const bf16x4_t*>(row + 16 * kb + 16);
      acc = mfma32(a0, a1, x[kb], x[kb + 1], acc);
    }
    if (NKB & 1) acc = mfma16(*reinterpret_cast<const bf16x4_t*>(row + 16 * (NKB - 1)), x[NKB - 1], acc);
    y[t] = acc;
  }
}

// dh[u] = W^T[16u .. 16u+15][:] . dz over NKB output blocks of W (transposed A
// operand read element-wise from the row-major weights).
template <int NU, int NKB>
__device__ __forceinline__ void dense_bwd(const uint16_t* W, int pitch, const bf16x4_t* dz, f32x4_t* dh, int lane) {
  const int r = lane & 15, g = lane >> 4;
  auto col = [&](int kb, int u) {
    const uint16_t* c = W + (16 * kb + 4 * g) * pitch + 16 * u + r;
    bf16x4_t a;
    a[0] = (short)c[0];
    a[1] = (short)c[pitch];
    a[2] = (short)c[2 * pitch];
    a[3] = (short)c[3 * pitch];
    return a;
  };
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb + 1 < NKB; kb += 2) acc = mfma32(col(kb, u), col(kb + 1, u), dz[kb], dz[kb + 1], acc);
    if (NKB & 1) acc = mfma16(col(NKB - 1, u), dz[NKB - 1], acc);
    dh[u] = acc;
  }
}

// dz = dh * (h > 0), cast to bf16; h holds post-ReLU bf16 activations.
__device__ __forceinline__ bf16x4_t relu_grad(f32x4_t dh, bf16x4_t h) {
  return pack_bf16(h[0] > 0 ? dh[0] : 0.0f, h[1] > 0 ? dh[1] : 0.0f, h[2] > 0 ? dh[2] : 0.0f,
                   h[3] > 0 ? dh[3] : 0.0f);
}

struct MlpInput {
  V3 p;  // position
  V3 wo; // direction towards the previous vertex
  V3 n;  // shading normal
  float rough;
};

// Network input of spatial-cache entry e (see mlp_cache_entry).
__device__ __forceinline__ MlpInput mlp_cache_input(const MlpView& m, uint32_t e) {
  const uint32_t res = m.cache_res, face = e % kMlpCacheFaces, cell = e / kMlpCacheFaces;
  const uint32_t x = cell % res, y = (cell / res) % res, z = cell / (res * res);
  const float size = 1.0f / m.inv_size, step = size / (float)res;
  V3 p{m.lo[0] + ((float)x + 0.5f) * step, m.lo[1] + ((float)y + 0.5f) * step, m.lo[2] + ((float)z + 0.5f) * step};
  float sgn = (face & 1u) ? -1.0f : 1.0f;
  V3 n{face / 2u == 0u ? sgn : 0.0f, face / 2u == 1u ? sgn : 0.0f, face / 2u == 2u ? sgn : 0.0f};
  return MlpInput{p, n, n, 0.0f};
}

struct MlpActs {
  bf16x4_t x0[3], h1[4], h2[4], h3[4];
};

// Input operand of one lane: levels {2g, 2g+1} (block 0), {8+2g, 9+2g}
// (block 1); block 2 holds wo, n, roughness and the constant 1.
__device__ __forceinline__ void mlp_inputs(const MlpView& m, const MlpInput& in, int lane, bf16x4_t* x0) {
  const int g = lane >> 4;
  float px = fminf(fmaxf((in.p.x - m.lo[0]) * m.inv_size, 0.0f), 1.0f);
  float py = fminf(fmaxf((in.p.y - m.lo[1]) * m.inv_size, 0.0f), 1.0f);
  float pz = fminf(fmaxf((in.p.z - m.lo[2]) * m.inv_size, 0.0f), 1.0f);
  float f[8];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
#pragma unroll
    for (int q = 0; q < 2; ++q) mlp_encode_level(m, 8 * b + 2 * g + q, px, py, pz, f[4 * b + 2 * q], f[4 * b + 2 * q + 1]);
  }
  x0[0] = pack_bf16(f[0], f[1], f[2], f[3]);
  x0[1] = pack_bf16(f[4], f[5], f[6], f[7]);
  if (g == 0) {
    x0[2] = pack_bf16(in.wo.x, in.wo.y, in.wo.z, in.n.x);
  } else if (g == 1) {
    x0[2] = pack_bf16(in.n.y, in.n.z, in.rough, 1.0f);
  } else {
    x0[2] = pack_bf16(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// Forward pass of one wave's 16 samples; o: the 48 raw outputs in D layout.
__device__ __forceinline__ void mlp_forward(const uint16_t* lw, MlpActs& a, f32x4_t* o, int lane) {
  f32x4_t y[4];
  dense_fwd<4, 3>(lw + kLdsW0, kLdsP0, a.x0, y, lane);
#pragma unroll
  for (int t = 0; t < 4; ++t) a.h1[t] = relu_bf16(y[t]);
  dense_fwd<4, 4>(lw + kLdsW1, kLdsP1, a.h1, y, lane);
#pragma unroll
  for (int t = 0; t < 4; ++t) a.h2[t] = relu_bf16(y[t]);
  dense_fwd<4, 4>(lw + kLdsW2, kLdsP1, a.h2, y, lane);
#pragma unroll
  for (int t = 0; t < 4; ++t) a.h3[t] = relu_bf16(y[t]);
  dense_fwd<3, 4>(lw + kLdsW3, kLdsP1, a.h3, o, lane);
}

// Mixture parameters of lobes 2g and 2g+1 of sample r from the raw outputs in
// the head scratch (row r); softmax over the 8 lobes across the four lanes
// that share the sample.
struct MlpHead {
  float pi[2], kappa[2], A[2], mnorm[2], sk[2];
  V3 mu[2];
  float c, sig;
};

__device__ __forceinline__ MlpHead mlp_head(const float* row, int lane) {
  const int g = lane >> 4;
  MlpHead h;
  float a[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const float* o = row + 5 * (2 * g + q);
    a[q] = o[0];
    V3 mv{o[1], o[2], o[3]};
    h.mnorm[q] = fmaxf(length(mv), 1e-12f);
    h.mu[q] = mv / h.mnorm[q];
    h.sk[q] = o[4];
    h.kappa[q] = expf(fminf(fmaxf(o[4], kMlpLogKappaMin), kMlpLogKappaMax));
  }
  float amax = fmaxf(a[0], a[1]);
  amax = fmaxf(amax, __shfl_xor(amax, 16));
  amax = fmaxf(amax, __shfl_xor(amax, 32));
  float e0 = expf(a[0] - amax), e1 = expf(a[1] - amax);
  float es = e0 + e1;
  es += __shfl_xor(es, 16);
  es += __shfl_xor(es, 32);
  h.pi[0] = e0 / es;
  h.pi[1] = e1 / es;
#pragma unroll
  for (int q = 0; q < 2; ++q) h.A[q] = h.pi[q] * vmf_norm(h.kappa[q]);
  h.sig = 1.0f / (1.0f + expf(-row[40]));
  h.c = kMlpSelMin + kMlpSelSpan * h.sig;
  return h;
}

// Spills the 48 raw outputs (D layout) into the head scratch rows.
__device__ __forceinline__ void mlp_spill(float* hs, const f32x4_t* o, int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) hs[r * kHeadPitch + 16 * t + 4 * g + j] = o[t][j];
}

// Writes the mixture of sample r in the renderer's per-vertex layout
// (kMlpMixF4 float4: 8 x (mu, kappa), pi[0..7], A[0..7]) plus c.
__device__ __forceinline__ void mlp_write_mix(const MlpHead& h, float* mx, float* cout, int lane) {
  const int g = lane >> 4;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int j = 2 * g + q;
    mx[4 * j + 0] = h.mu[q].x;
    mx[4 * j + 1] = h.mu[q].y;
    mx[4 * j + 2] = h.mu[q].z;
    mx[4 * j + 3] = h.kappa[q];
    mx[32 + j] = h.pi[q];
    mx[40 + j] = h.A[q];
  }
  if (g == 0) *cout = h.c;
}

}  // namespace rtpg
