// Learned directional guide: a multiresolution hash-grid encoding of the
// shading position, concatenated with the outgoing direction, the normal and
// the roughness, feeds a 3 x 64 ReLU MLP whose 41 outputs parameterise an
// 8-lobe von Mises-Fisher mixture and the probability c of sampling that
// mixture instead of the BSDF.  No counterpart in the reference.
//
// Layout of one forward pass on the matrix cores: a wave evaluates 16 samples.
// Weights are the A operand (rows = output features) and activations the B
// operand (columns = samples) of v_mfma_f32_16x16x16_bf16, so the D tile of
// one layer (lane l: sample l%16, features 4*(l/16)+j) is already the B
// operand of the next layer after ReLU and the cast to bf16.  Weights are
// bf16 in LDS; accumulation is fp32.
#pragma once
#include "common.h"

namespace rtpg {

constexpr int kMlpLevels = 16;
constexpr int kMlpFeat = 2;
constexpr int kMlpHidden = 64;
constexpr int kMlpLobes = 8;
constexpr int kMlpIn = 48;       // 32 encoded + wo(3) + n(3) + roughness + 1 + 8 zeros
constexpr int kMlpOut = 48;      // 8 x (logit, m xyz, log kappa) + selection logit + 7 zeros
constexpr int kMlpOutUsed = 41;
constexpr int kMlpW0 = 0;                                   // 64 x 48 (out x in, row-major)
constexpr int kMlpW1 = kMlpW0 + kMlpHidden * kMlpIn;        // 64 x 64
constexpr int kMlpW2 = kMlpW1 + kMlpHidden * kMlpHidden;    // 64 x 64
constexpr int kMlpW3 = kMlpW2 + kMlpHidden * kMlpHidden;    // 48 x 64
constexpr int kMlpWeights = kMlpW3 + kMlpOut * kMlpHidden;  // 14336
constexpr float kMlpLogKappaMin = -3.0f;
constexpr float kMlpLogKappaMax = 8.0f;
constexpr float kMlpSelMin = 0.05f;   // c = kMlpSelMin + kMlpSelSpan * sigmoid(t)
constexpr float kMlpSelSpan = 0.9f;
constexpr int kMlpMixF4 = 12;         // per-vertex mixture: 8 x (mu, kappa), pi[8], A[8]
// Levels kMlpPkFirst.. (hashed at the default 2^19 entries: a few
// contributions per entry and training step) accumulate their training
// gradient as packed bf16 pairs, one atomic per entry for both features; the
// dense coarse levels, where many samples share an entry, in fp32.
constexpr int kMlpPkFirst = 5;

struct MlpView {
  const uint32_t* table;  // [levels][T]: two bf16 features per entry
  float* grad;            // [levels][T][2] fp32 gradient (training)
  uint32_t* grad_pk;      // [levels - kMlpPkFirst][T] bf16 pair gradient of the fine levels (training)
  const uint16_t* w;      // kMlpWeights bf16
  float lo[3];
  float inv_size;         // position -> [0,1]^3 of the cubic box
  uint32_t tmask;         // T - 1
  uint32_t dense;         // bit l: level l indexed densely
  float res[kMlpLevels];  // grid resolution of level l (integral)
  uint32_t ready;         // trained: the renderer may sample from it
  uint32_t cache_res;     // spatial cache: cells per axis of the box (0: no cache)
};

// Spatial cache of the guide: one mixture per (cell of a cache_res^3 grid
// over the network's cubic box, normal direction), the network evaluated at
// the cell centre with that normal, wo = normal and roughness 0.  Entry =
// cell * 6 + face, face 0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z (largest normal
// component).  The renderer samples and evaluates the cached mixture of the
// vertex's entry, so proposal and density stay consistent.
constexpr uint32_t kMlpCacheFaces = 6;

RTPG_HD uint32_t mlp_cache_face(V3 n) {
  float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
  if (ax >= ay && ax >= az) return n.x >= 0.0f ? 0u : 1u;
  if (ay >= az) return n.y >= 0.0f ? 2u : 3u;
  return n.z >= 0.0f ? 4u : 5u;
}

RTPG_HD uint32_t mlp_cache_cell1(float p, float lo, float inv_size, uint32_t res) {
  float q = (p - lo) * inv_size * (float)res;
  q = fminf(fmaxf(floorf(q), 0.0f), (float)(res - 1u));
  return (uint32_t)q;
}

RTPG_HD uint32_t mlp_cache_entry(const float* lo, float inv_size, uint32_t res, V3 p, V3 n) {
  uint32_t x = mlp_cache_cell1(p.x, lo[0], inv_size, res);
  uint32_t y = mlp_cache_cell1(p.y, lo[1], inv_size, res);
  uint32_t z = mlp_cache_cell1(p.z, lo[2], inv_size, res);
  return ((z * res + y) * res + x) * kMlpCacheFaces + mlp_cache_face(n);
}

// ------------------------------------------------------------- bf16 helpers
RTPG_HD uint16_t f2bf(float f) {
  uint32_t u = fbits(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0u;
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
RTPG_HD float bf2f(uint16_t h) { return bitsf((uint32_t)h << 16); }

// ------------------------------------------------------------- hash encoding
// Corner index of level l (instant-NGP spatial hash, or dense for coarse levels).
RTPG_HD uint32_t mlp_index(const MlpView& m, int l, uint32_t x, uint32_t y, uint32_t z) {
  if ((m.dense >> l) & 1u) {
    uint32_t r1 = (uint32_t)m.res[l] + 1u;
    return x + r1 * (y + r1 * z);
  }
  return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & m.tmask;
}

struct MlpCell {
  uint32_t x, y, z;
  float fx, fy, fz;
};

RTPG_HD MlpCell mlp_cell(const MlpView& m, int l, float px, float py, float pz) {
  float r = m.res[l];
  float gx = px * r, gy = py * r, gz = pz * r;
  float ix = fminf(fmaxf(floorf(gx), 0.0f), r - 1.0f);
  float iy = fminf(fmaxf(floorf(gy), 0.0f), r - 1.0f);
  float iz = fminf(fmaxf(floorf(gz), 0.0f), r - 1.0f);
  MlpCell c;
  c.x = (uint32_t)ix;
  c.y = (uint32_t)iy;
  c.z = (uint32_t)iz;
  c.fx = fminf(fmaxf(gx - ix, 0.0f), 1.0f);
  c.fy = fminf(fmaxf(gy - iy, 0.0f), 1.0f);
  c.fz = fminf(fmaxf(gz - iz, 0.0f), 1.0f);
  return c;
}

// Trilinear interpolation of the two features of level l at p in [0,1]^3.
__device__ __forceinline__ void mlp_encode_level(const MlpView& m, int l, float px, float py, float pz, float& f0,
                                                 float& f1) {
  MlpCell c = mlp_cell(m, l, px, py, pz);
  const uint32_t* tab = m.table + (size_t)l * (m.tmask + 1u);
  f0 = 0.0f;
  f1 = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    float w = (dx ? c.fx : 1.0f - c.fx) * (dy ? c.fy : 1.0f - c.fy) * (dz ? c.fz : 1.0f - c.fz);
    uint32_t e = tab[mlp_index(m, l, c.x + dx, c.y + dy, c.z + dz)];
    f0 += w * bf2f((uint16_t)(e & 0xFFFFu));
    f1 += w * bf2f((uint16_t)(e >> 16));
  }
}

// Scatter d loss / d (features of level l) to the fp32 gradient table.
__device__ __forceinline__ void mlp_encode_level_grad(const MlpView& m, int l, float px, float py, float pz, float g0,
                                                      float g1) {
  MlpCell c = mlp_cell(m, l, px, py, pz);
  float* gt = m.grad + (size_t)l * (m.tmask + 1u) * 2u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    float w = (dx ? c.fx : 1.0f - c.fx) * (dy ? c.fy : 1.0f - c.fy) * (dz ? c.fz : 1.0f - c.fz);
    uint32_t i = mlp_index(m, l, c.x + dx, c.y + dy, c.z + dz);
    if (g0 != 0.0f) unsafeAtomicAdd(gt + 2u * i, w * g0);
    if (g1 != 0.0f) unsafeAtomicAdd(gt + 2u * i + 1u, w * g1);
  }
}

// The same scatter for a fine level l >= kMlpPkFirst into the packed bf16
// gradient: both features of a corner in one global_atomic_pk_add_bf16 (the
// contributions are rounded to bf16, the sum is kept in bf16; the fine levels'
// entries receive a few contributions per training step).
typedef __bf16 bf16x2_v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void mlp_encode_level_grad_pk(const MlpView& m, int l, float px, float py, float pz, float g0,
                                                         float g1) {
  MlpCell c = mlp_cell(m, l, px, py, pz);
  bf16x2_v* gt = reinterpret_cast<bf16x2_v*>(m.grad_pk + (size_t)(l - kMlpPkFirst) * (m.tmask + 1u));
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    float w = (dx ? c.fx : 1.0f - c.fx) * (dy ? c.fy : 1.0f - c.fy) * (dz ? c.fz : 1.0f - c.fz);
    uint32_t i = mlp_index(m, l, c.x + dx, c.y + dy, c.z + dz);
    float a = w * g0, b = w * g1;
    if (a != 0.0f || b != 0.0f) {
      bf16x2_v v = {(__bf16)a, (__bf16)b};
      __builtin_amdgcn_global_atomic_fadd_v2bf16(gt + i, v);
    }
  }
}

// Work-group table of gradient sums in LDS: key level x entries + entry (the
// entry's pair in the fp32 gradient), two fp32 features per slot, 16 linear
// probes; the training kernel flushes it to the fp32 gradient at its end.  A work-group's samples are a
// contiguous slice of a spatially ordered window, so they touch few coarse
// entries; a full table makes the caller add to global memory directly.
constexpr uint32_t kMlpGradTab = 1024;
constexpr uint32_t kMlpEmptyKey = 0xFFFFFFFFu;
struct MlpGradTable {
  uint32_t* keys;  // kMlpGradTab, or NULL: no table
  float* vals;     // 2 x kMlpGradTab
};

__device__ __forceinline__ bool mlp_tab_add(const MlpGradTable& t, uint32_t key, float a, float b) {
  uint32_t h = (key * 0x9E3779B1u) & (kMlpGradTab - 1u);
  for (int probe = 0; probe < 16; ++probe) {
    uint32_t prev = atomicCAS(&t.keys[h], kMlpEmptyKey, key);
    if (prev == kMlpEmptyKey || prev == key) {
      if (a != 0.0f) atomicAdd(&t.vals[2u * h], a);
      if (b != 0.0f) atomicAdd(&t.vals[2u * h + 1u], b);
      return true;
    }
    h = (h + 1u) & (kMlpGradTab - 1u);
  }
  return false;
}

// The same scatter for levels where neighbouring samples often share a grid
// cell.  Each row of 16 lanes holds 16 samples at one level (lane r of the
// row: sample r; every lane of the row takes part, valid = false lanes
// contribute nothing).  Runs of consecutive lanes in the same cell sum their
// 8 corner x 2 feature contributions with a segmented suffix scan and the
// first lane of a run issues one atomic per corner (kPacked: one packed bf16
// pair into grad_pk, for levels >= kMlpPkFirst; else one fp32 add per
// feature, into the work-group table tab when one is given) for the whole
// run.  Needs res[l] < 1024 (the cell key packs 10 bits per axis).
template <bool kPacked>
__device__ __forceinline__ void mlp_encode_level_grad_runs(const MlpView& m, int l, float px, float py, float pz,
                                                           float g0, float g1, bool valid,
                                                           const MlpGradTable& tab = MlpGradTable{nullptr, nullptr}) {
  MlpCell c = mlp_cell(m, l, px, py, pz);
  const uint32_t lane = __lane_id(), r = lane & 15u;
  const uint32_t key = valid ? (c.x | (c.y << 10) | (c.z << 20)) : 0xFFFFFFFFu;
  const uint32_t kn = (uint32_t)__shfl_down((int)key, 1, 16);
  const uint32_t kp = (uint32_t)__shfl_up((int)key, 1, 16);
  const bool same_next = r != 15u && kn == key;
  float v[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    float w = (dx ? c.fx : 1.0f - c.fx) * (dy ? c.fy : 1.0f - c.fy) * (dz ? c.fz : 1.0f - c.fz);
    v[2 * k] = valid ? w * g0 : 0.0f;
    v[2 * k + 1] = valid ? w * g1 : 0.0f;
  }
  if (__ballot(same_next) != 0ull) {
    // last lane of this lane's run: the first run end at or after it (lane 15
    // of every row is one)
    const uint64_t ends = __ballot(!same_next);
    const uint32_t e = lane + (uint32_t)__ffsll((long long)(ends >> lane)) - 1u;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        float t = __shfl_down(v[j], off, 16);
        if (lane + (uint32_t)off <= e) v[j] += t;
      }
    }
  }
  if (valid && (r == 0u || kp != key)) {
    if (kPacked) {
      bf16x2_v* gt = reinterpret_cast<bf16x2_v*>(m.grad_pk + (size_t)(l - kMlpPkFirst) * (m.tmask + 1u));
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        uint32_t i = mlp_index(m, l, c.x + dx, c.y + dy, c.z + dz);
        if (v[2 * k] == 0.0f && v[2 * k + 1] == 0.0f) continue;
        if (tab.keys && mlp_tab_add(tab, (uint32_t)l * (m.tmask + 1u) + i, v[2 * k], v[2 * k + 1])) continue;
        bf16x2_v pv = {(__bf16)v[2 * k], (__bf16)v[2 * k + 1]};
        __builtin_amdgcn_global_atomic_fadd_v2bf16(gt + i, pv);
      }
    } else {
      float* gt = m.grad + (size_t)l * (m.tmask + 1u) * 2u;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        uint32_t i = mlp_index(m, l, c.x + dx, c.y + dy, c.z + dz);
        if (v[2 * k] == 0.0f && v[2 * k + 1] == 0.0f) continue;
        if (tab.keys && mlp_tab_add(tab, (uint32_t)l * (m.tmask + 1u) + i, v[2 * k], v[2 * k + 1])) continue;
        if (v[2 * k] != 0.0f) unsafeAtomicAdd(gt + 2u * i, v[2 * k]);
        if (v[2 * k + 1] != 0.0f) unsafeAtomicAdd(gt + 2u * i + 1u, v[2 * k + 1]);
      }
    }
  }
}

// ------------------------------------------------------------- mixture
constexpr float kLog2E = 1.44269504088896340736f;
// Unit-sphere vMF: v(w) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)).
RTPG_HD float vmf_norm(float kappa) { return kappa / (2.0f * kPi * -expm1f(-2.0f * kappa)); }

// Mixture density at unit direction d; mx: kMlpMixF4 float4 of one vertex.
// The exponentials use the hardware base-2 exponential (v_exp_f32, 1 ulp):
// exp(x) = 2^(x log2 e); the arguments are <= 0, so no overflow path is needed.
__device__ __forceinline__ float mix_pdf(const float4* mx, V3 d) {
  float4 a0 = mx[10], a1 = mx[11];
  float A[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  float q = 0.0f;
#pragma unroll
  for (int j = 0; j < kMlpLobes; ++j) {
    float4 l = mx[j];
    float x = l.w * (l.x * d.x + l.y * d.y + l.z * d.z - 1.0f);
    q += A[j] * __builtin_amdgcn_exp2f(x * kLog2E);
  }
  return q;
}

// vMF sample around mu (W. Jakob, "Numerically stable sampling of the von
// Mises Fisher distribution on S^2", 2012).
RTPG_HD V3 vmf_sample(V3 mu, float kappa, float u1, float u2) {
  float a = fmaxf(u1 + (1.0f - u1) * expf(-2.0f * kappa), 1e-38f);
  float w = fminf(fmaxf(1.0f + logf(a) / kappa, -1.0f), 1.0f);
  float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  float phi = 2.0f * kPi * u2;
  Onb o = make_onb(mu);
  return normalize(o.t * (s * cosf(phi)) + o.b * (s * sinf(phi)) + o.n * w);
}

__device__ __forceinline__ V3 mix_sample(const float4* mx, Pcg32& rng) {
  float4 p0 = mx[8], p1 = mx[9];
  float P[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
  float u = rnd(rng);
  int j = kMlpLobes - 1;
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < kMlpLobes - 1; ++k) {
    acc += P[k];
    if (u < acc) {
      j = k;
      break;
    }
  }
  float4 l = mx[0];
#pragma unroll
  for (int k = 1; k < kMlpLobes; ++k)
    if (k == j) l = mx[k];
  float u1 = rnd(rng), u2 = rnd(rng);
  // vmf_sample with the azimuth's sine and cosine from the hardware
  // (v_sin_f32 / v_cos_f32 take the angle in revolutions: sin(2 pi u2))
  float kappa = l.w;
  float a = fmaxf(u1 + (1.0f - u1) * expf(-2.0f * kappa), 1e-38f);
  float w = fminf(fmaxf(1.0f + logf(a) / kappa, -1.0f), 1.0f);
  float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  Onb o = make_onb(V3{l.x, l.y, l.z});
  return normalize(o.t * (s * __builtin_amdgcn_cosf(u2)) + o.b * (s * __builtin_amdgcn_sinf(u2)) + o.n * w);
}

}  // namespace rtpg
