// Wavefront path tracer for gfx950: camera -> (extend -> shade -> shadow) x depth
// -> accumulate, plus the guide-record kernels.
//
// Replaces the OptiX raygen/closest-hit/miss programs of the reference
// (SimpleTrace.cu:45-103, 214-425, 625-630; SimpleGuide.cu:48-156, 287-511;
// SimpleGuide2.cu:48-153, 285-523).  Paths live in structure-of-arrays HBM
// buffers; between bounces the surviving paths are compacted into a queue with
// one atomic per work-group (ballot + popcount, combined in LDS) on the
// counter of the queue region that holds the work-group's slots, so the work
// of each bounce is the live paths only; the grid of every bounce is sized
// for the whole chunk and work-groups past their region's live count exit at
// once.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "guide_dev.h"
#include "mlp_tile.h"
#include "render.h"
#include "views.h"

namespace rtpg {

constexpr uint32_t kBlock = 128;

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

typedef _Float16 half4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ V3 load3(float4 v) { return V3{v.x, v.y, v.z}; }

// The queue region of a work-group whose first slot is b0: its first slot
// (base) and its live entries at this depth in counter array cnt.
struct QueueRegion {
  uint32_t base, live, index;
};
__device__ __forceinline__ QueueRegion queue_region(const PathSoA& ps, const uint32_t* cnt, uint32_t depth,
                                                    uint32_t b0) {
  QueueRegion q;
  q.index = b0 / ps.qreg;
  q.base = q.index * ps.qreg;
  q.live = cnt[depth * kQueueRegions + q.index];
  return q;
}

// Paths whose guide vertices are recorded: pid < rec_limit, or with the
// learned guide (mlp_stride > 0) every mlp_stride-th path from mlp_phase
// (training slot pid / mlp_stride).
__device__ __forceinline__ bool records_path(const PathSoA& ps, uint32_t pid) {
  return ps.mlp_stride ? pid % ps.mlp_stride == ps.mlp_phase : pid < ps.rec_limit;
}

// Moller-Trumbore, two-sided, in a form the CPU oracle repeats operation for
// operation, so hits are bit-identical: the cross and dot products fuse their
// last products (fmaf rounds once, on both sides alike), the barycentric
// tests compare the unnormalised u det and v det against det, and the one
// division is made only for a triangle that passes them.
__device__ __forceinline__ bool isect_tri(const TriRec* tp, V3 o, V3 d, float tmin, float& tbest,
                                          float& ou, float& ov) {
#pragma clang fp contract(off)
  const float4* q = reinterpret_cast<const float4*>(tp);
  float4 a = q[0], b = q[1], c = q[2];
  float e1x = b.x, e1y = b.y, e1z = b.z;
  float e2x = c.x, e2y = c.y, e2z = c.z;
  float px = fmaf(d.y, e2z, -(d.z * e2y));
  float py = fmaf(d.z, e2x, -(d.x * e2z));
  float pz = fmaf(d.x, e2y, -(d.y * e2x));
  float det = fmaf(e1x, px, fmaf(e1y, py, e1z * pz));
  if (det == 0.0f) return false;
  float sx = o.x - a.x, sy = o.y - a.y, sz = o.z - a.z;
  float ud = fmaf(sx, px, fmaf(sy, py, sz * pz));
  // u = ud / det in [0, 1]
  if (det > 0.0f ? !(ud >= 0.0f && ud <= det) : !(ud <= 0.0f && ud >= det)) return false;
  float qx = fmaf(sy, e1z, -(sz * e1y));
  float qy = fmaf(sz, e1x, -(sx * e1z));
  float qz = fmaf(sx, e1y, -(sy * e1x));
  float vd = fmaf(d.x, qx, fmaf(d.y, qy, d.z * qz));
  float uvd = ud + vd;
  // v >= 0 and u + v <= 1
  if (det > 0.0f ? !(vd >= 0.0f && uvd <= det) : !(vd <= 0.0f && uvd >= det)) return false;
  float inv = 1.0f / det;
  float t = fmaf(e2x, qx, fmaf(e2y, qy, e2z * qz)) * inv;
  if (!(t > tmin && t < tbest)) return false;
  tbest = t;
  ou = ud * inv;
  ov = vd * inv;
  return true;
}

// Traversal stack entries: (node or leaf, entry distance of its box), the
// node or leaf word as the BVH stores it (kLeafEntry in bvh.h).
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;  // never a node (< 2^31) nor a leaf (first < 2^27)

// A pointer into LDS (address space 3): accesses through it compile to
// ds_read / ds_write instead of flat loads and stores, which take the slower
// vector-memory path.  An entry is one 64-bit word (low: entry, high: bits
// of its distance).
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

struct StackRef {
  lds_u64* lds;      // this lane's column of the LDS short stack
  uint2* deep;       // this ray's column of the spill stack (stride = deep_stride)
  uint32_t deep_stride;
  uint32_t cap;      // total capacity
};

// The spill side uses nontemporal accesses: besides streaming the rarely
// used spill entries, this keeps the compiler from merging the two sides into
// one access through a selected generic pointer (a flat load or store, which
// takes the slower vector-memory path even for the LDS entries).
__device__ __forceinline__ void stack_push(const StackRef& s, uint32_t& sp, uint32_t e, float t) {
  uint2 v = make_uint2(e, fbits(t));
  if (sp < kShortStack) {
    s.lds[sp * kBlock] = ((unsigned long long)v.y << 32) | v.x;
  } else {
    uint2* d = s.deep + (size_t)(sp - kShortStack) * s.deep_stride;
    __builtin_nontemporal_store(v.x, &d->x);
    __builtin_nontemporal_store(v.y, &d->y);
  }
  ++sp;
}
__device__ __forceinline__ uint2 stack_pop(const StackRef& s, uint32_t& sp) {
  --sp;
  if (sp < kShortStack) {
    unsigned long long w = s.lds[sp * kBlock];
    return make_uint2((uint32_t)w, (uint32_t)(w >> 32));
  }
  const uint2* d = s.deep + (size_t)(sp - kShortStack) * s.deep_stride;
  return make_uint2(__builtin_nontemporal_load(&d->x), __builtin_nontemporal_load(&d->y));
}

#define RTPG_CS(a, b)                         \
  if (tk[b] < tk[a]) {                        \
    float tt = tk[a]; tk[a] = tk[b]; tk[b] = tt; \
    uint32_t nn = nk[a]; nk[a] = nk[b]; nk[b] = nn; \
  }

// BVH4 traversal: closest hit (kAny = false) or any hit (kAny = true).
// Two phases per round, so that a wave tests triangles together: the inner
// phase visits nodes nearest-first until the lane's next entry is a leaf (or
// nothing is left), pushing the other hit children, inner nodes and leaves
// alike, far-to-near with their entry distance; the leaf phase tests that
// leaf's triangles.  A popped entry whose box is entered beyond the best hit
// (with the same 2-ulp slack as the box test) is dropped without a fetch.
// kCount adds the nodes fetched and triangles tested to wcnt[0], wcnt[1]
// (the algorithmic-byte accounting of the roofline).
template <bool kAny, bool kCount = false>
__device__ bool traverse(const SceneView& sc, V3 o, V3 d, float tmin, float tmax, const StackRef& st,
                         uint32_t* err, float& t_out, float& u_out, float& v_out, int32_t& tri_out,
                         uint32_t* wcnt = nullptr) {
  const float eps = 1e-20f;
  V3 dd = V3{fabsf(d.x) > eps ? d.x : copysignf(eps, d.x), fabsf(d.y) > eps ? d.y : copysignf(eps, d.y),
             fabsf(d.z) > eps ? d.z : copysignf(eps, d.z)};
  V3 inv = V3{1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z};
  V3 oi = o * inv;
  // per axis, whether the ray enters a box through its upper plane
  const bool rx = inv.x < 0.0f, ry = inv.y < 0.0f, rz = inv.z < 0.0f;
  float tbest = tmax;
  int32_t best = -1;
  float bu = 0.0f, bv = 0.0f;
  uint32_t sp = 0;
  // cur: the next entry (inner node, leaf, or kNoEntry) and its entry
  // distance; pleaf: a leaf met in the inner phase and postponed
  uint32_t cur = 0;  // the root
  float cur_t = 0.0f;
  uint32_t pleaf = kNoEntry;
  // next stacked entry whose box can still hold a closer hit
  auto pop_valid = [&]() {
    cur = kNoEntry;
    while (sp > 0) {
      uint2 e = stack_pop(st, sp);
      if (bitsf(e.y) <= tbest * 1.00000024f) {
        cur = e.x;
        cur_t = bitsf(e.y);
        break;
      }
    }
  };
  for (;;) {
    // ---- inner phase (speculative): a lane that meets a leaf keeps it aside
    // and goes on visiting inner nodes until every lane of the wave holds a
    // leaf or has nothing left, so that the leaf phase runs with full waves
    while (cur != kNoEntry && !(cur & kLeafEntry)) {
      if (kCount) ++wcnt[0];
      // the planes are binary16 offsets from the node's corner (bvh.h): the
      // corner's ray distance once per node, then one v_fma_mix_f32 per plane
      // Entry and exit planes per axis are the lower and upper ones, swapped
      // where the ray runs downwards (fma is monotonic in the plane, so this
      // equals a min and a max of the two distances).
      const uint4* np = reinterpret_cast<const uint4*>(sc.nodes + cur);
      const uint4 qx = np[0], qy = np[1], qz = np[2];
      const uint4 ch = np[3];
      const float4 org = reinterpret_cast<const float4*>(np)[4];
      const half4 nx = __builtin_bit_cast(half4, rx ? make_uint2(qx.z, qx.w) : make_uint2(qx.x, qx.y));
      const half4 fx = __builtin_bit_cast(half4, rx ? make_uint2(qx.x, qx.y) : make_uint2(qx.z, qx.w));
      const half4 ny = __builtin_bit_cast(half4, ry ? make_uint2(qy.z, qy.w) : make_uint2(qy.x, qy.y));
      const half4 fy = __builtin_bit_cast(half4, ry ? make_uint2(qy.x, qy.y) : make_uint2(qy.z, qy.w));
      const half4 nz = __builtin_bit_cast(half4, rz ? make_uint2(qz.z, qz.w) : make_uint2(qz.x, qz.y));
      const half4 fz = __builtin_bit_cast(half4, rz ? make_uint2(qz.x, qz.y) : make_uint2(qz.z, qz.w));
      const float cx = fmaf(org.x, inv.x, -oi.x), cy = fmaf(org.y, inv.y, -oi.y), cz = fmaf(org.z, inv.z, -oi.z);
      float tk[4];
      uint32_t nk[4];
      int n_hit = 0;
      uint32_t hmask = 0;  // children whose box the ray enters
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float t0x = fmaf((float)nx[i], inv.x, cx), t1x = fmaf((float)fx[i], inv.x, cx);
        float t0y = fmaf((float)ny[i], inv.y, cy), t1y = fmaf((float)fy[i], inv.y, cy);
        float t0z = fmaf((float)nz[i], inv.z, cz), t1z = fmaf((float)fz[i], inv.z, cz);
        float tn = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
        float tf = fminf(fminf(t1x, t1y), fminf(t1z, tbest));
        bool hit = tn <= tf * 1.00000024f;
        tk[i] = hit ? tn : 3.0e38f;
        nk[i] = i == 0 ? ch.x : (i == 1 ? ch.y : (i == 2 ? ch.z : ch.w));
        n_hit += hit ? 1 : 0;
        hmask |= hit ? 1u << i : 0u;
      }
      if (n_hit == 0) {
        pop_valid();
      } else {
        // Closest hit visits the children nearest first (a sorting network
        // on the entry distances) when some lane of the wave entered more
        // than one; any hit needs no order, as every entered box is visited
        // or culled against the same tmax until a hit ends it.
        const bool sorted = !kAny && __any(n_hit > 1);
        if (sorted) {
          RTPG_CS(0, 1) RTPG_CS(2, 3) RTPG_CS(0, 2) RTPG_CS(1, 3) RTPG_CS(1, 2)
        }
        if (sp + (uint32_t)(n_hit - 1) > st.cap) {
          atomicOr(err, 1u);  // stack bound of the build violated: stop this ray
          cur = kNoEntry;
          pleaf = kNoEntry;
          sp = 0;
        } else if (sorted) {
          cur = nk[0];
          cur_t = tk[0];
          // the other hit children, far to near (static indices: no
          // register-array selects)
#pragma unroll
          for (int k = 3; k >= 1; --k)
            if (k < n_hit) stack_push(st, sp, nk[k], tk[k]);
        } else {
          // the lowest entered child goes on, the others are stacked
          const int first = __builtin_ctz(hmask);
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            if (k == first) {
              cur = nk[k];
              cur_t = tk[k];
            } else if (hmask & (1u << k)) {
              stack_push(st, sp, nk[k], tk[k]);
            }
          }
        }
      }
      if (cur != kNoEntry && (cur & kLeafEntry) && pleaf == kNoEntry) {
        pleaf = cur;
        pop_valid();
      }
      if (__all(pleaf != kNoEntry || cur == kNoEntry || (cur & kLeafEntry))) break;
    }
    // ---- leaf phase: the postponed leaf, else the current entry if a leaf
    uint32_t leaf;
    if (pleaf != kNoEntry) {
      leaf = pleaf;
      pleaf = kNoEntry;
    } else if (cur != kNoEntry && (cur & kLeafEntry)) {
      leaf = cur;
      cur = kNoEntry;
    } else if (cur == kNoEntry) {
      break;
    } else {
      continue;  // an inner node left by the wave's early exit
    }
    uint32_t first = leaf & 0x07FFFFFFu, cnt = ((leaf >> 27) & 0xFu) + 1u;
    for (uint32_t k = 0; k < cnt; ++k) {
      if (kCount) ++wcnt[1];
      if (isect_tri(sc.tris + first + k, o, d, tmin, tbest, bu, bv)) {
        best = (int32_t)(first + k);
        if (kAny) {
          t_out = tbest;
          u_out = bu;
          v_out = bv;
          tri_out = best;
          return true;
        }
      }
    }
    // the next entry, unless it is still pending and can hold a closer hit
    if (cur == kNoEntry || cur_t > tbest * 1.00000024f) pop_valid();
  }
  t_out = tbest;
  u_out = bu;
  v_out = bv;
  tri_out = best;
  return best >= 0;
}
#undef RTPG_CS

__device__ __forceinline__ StackRef make_stack(uint2 (*lds)[kBlock], uint32_t* deep, uint32_t deep_stride,
                                               uint32_t slot, uint32_t cap) {
  StackRef s;
  s.lds = (lds_u64*)&lds[0][threadIdx.x];
  s.deep = deep ? reinterpret_cast<uint2*>(deep) + slot : nullptr;
  s.deep_stride = deep_stride;
  s.cap = cap;
  return s;
}

// ------------------------------------------------------------------ ray queries (tests, shadow)
__global__ void __launch_bounds__(kBlock) k_trace_rays(SceneView sc, const float4* rays, uint32_t n,
                                                       float4* hits, uint32_t* occl, uint32_t any,
                                                       uint32_t* deep, uint32_t* err) {
  __shared__ uint2 lds[kShortStack][kBlock];
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 a = rays[2 * i], b = rays[2 * i + 1];
  StackRef st = make_stack(lds, deep, n, i, sc.max_stack);
  float t, u, v;
  int32_t tri;
  if (any) {
    bool h = traverse<true>(sc, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, a.w, b.w, st, err, t, u, v, tri);
    occl[i] = h ? 1u : 0u;
  } else {
    traverse<false>(sc, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, a.w, b.w, st, err, t, u, v, tri);
    uint32_t prim = tri >= 0 ? fbits(sc.tris[tri].v0[3]) : 0xFFFFFFFFu;
    reinterpret_cast<uint4*>(hits)[i] = make_uint4(fbits(tri >= 0 ? t : INFINITY), fbits(u), fbits(v), prim);
  }
}

// ------------------------------------------------------------------ wavefront kernels
__global__ void __launch_bounds__(kBlock) k_camera(RenderParams rp, Camera cam, PathSoA ps) {
  uint32_t pid = blockIdx.x * kBlock + threadIdx.x;
  if (pid >= ps.n) return;
  uint32_t npix = rp.width * rp.height;
  uint32_t pix = pid % npix;
  uint32_t s = pid / npix;
  uint32_t x = pix % rp.width, y = pix / rp.width;
  Pcg32 rng = path_rng(rp.seed, pix, rp.sample_offset + s);
  float jx = rnd(rng), jy = rnd(rng);
  float dx = 2.0f * ((float)x + jx) / (float)rp.width - 1.0f;
  float dy = 2.0f * ((float)y + jy) / (float)rp.height - 1.0f;
  V3 U{cam.u[0], cam.u[1], cam.u[2]}, Vv{cam.v[0], cam.v[1], cam.v[2]}, W{cam.w[0], cam.w[1], cam.w[2]};
  V3 d = normalize(dx * U + dy * Vv + W);
  // depth 0: the path's slot is its index
  ps.ray_o[0][pid] = make_float4(cam.eye[0], cam.eye[1], cam.eye[2], 0.0f);
  ps.ray_d[0][pid] = make_float4(d.x, d.y, d.z, bitsf(kMetaDelta));  // depth 0, counts emission
  ps.T[0][pid] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  ps.L[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ps.rng[0][pid] = make_uint2((uint32_t)rng.state, (uint32_t)(rng.state >> 32));
  ps.queue[0][pid] = pid;
}

// Adds per-lane work counters into 64-bit totals, one atomic per wave.
__device__ __forceinline__ void wave_count(unsigned long long* dst, uint32_t a, uint32_t b) {
  unsigned long long x = a, y = b;
  for (int off = 32; off >= 1; off >>= 1) {
    x += __shfl_xor(x, off);
    y += __shfl_xor(y, off);
  }
  if (lane_id() == 0) {
    atomicAdd(dst, x);
    atomicAdd(dst + 1, y);
  }
}

// 8 waves per SIMD (64 VGPRs): traversal hides its dependent node fetches
// with occupancy.
template <bool kCount>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_extend(SceneView sc, PathSoA ps, uint32_t depth) {
  __shared__ uint2 lds[kShortStack][kBlock];
  const uint32_t b0 = blockIdx.x * kBlock, i = b0 + threadIdx.x;
  const QueueRegion q = queue_region(ps, ps.count, depth, b0);
  if (b0 - q.base >= q.live) return;  // uniform over the work-group
  uint32_t cnt[2] = {0u, 0u};
  if (i - q.base < q.live) {
    float4 a = ps.ray_o[depth & 1][i], b = ps.ray_d[depth & 1][i];
    StackRef st = make_stack(lds, ps.deep, ps.n, i, sc.max_stack);
    float t, u, v;
    int32_t tri;
    traverse<false, kCount>(sc, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, kRayTmin, kRayTmax, st, ps.err, t, u, v, tri,
                            cnt);
    ps.hit[i] = make_float4(t, u, v, bitsf((uint32_t)tri));
  }
  if (kCount) wave_count(ps.work, cnt[0], cnt[1]);
}

__device__ __forceinline__ uint32_t pick_light(const SceneView& sc, float u) {
  uint32_t lo = 0, hi = sc.n_lights - 1u;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (sc.light_cdf[mid] > u) hi = mid; else lo = mid + 1u;
  }
  return lo;
}

struct LightSample {
  V3 p, n, Le;
  float area, pdf_sel;
  uint32_t idx;
};

// Uniform point on a light triangle, barycentrics of Random.h:290-296.
template <bool kTex>
__device__ __forceinline__ LightSample sample_light(const SceneView& sc, Pcg32& rng) {
  float ul = rnd(rng);
  float u1 = rnd(rng), u2 = rnd(rng);
  uint32_t li = pick_light(sc, ul);
  const float4* L = sc.lights + 4u * li;
  float4 a = L[0], b = L[1], c = L[2], e = L[3];
  float lo = fminf(u1, u2), hi = fmaxf(u1, u2);
  LightSample s;
  V3 v0{a.x, a.y, a.z}, e1{b.x, b.y, b.z}, e2{c.x, c.y, c.z};
  s.p = v0 + e1 * (1.0f - hi) + e2 * (hi - lo);
  s.n = normalize(cross(e1, e2));
  s.Le = V3{e.x, e.y, e.z};
  if (kTex && sc.tattr) {
    // MeshLight::GetEmission at the sampled point's texture coordinate
    // (SimpleKernel.h:81-104): barycentrics (lo, 1 - hi, hi - lo)
    uint32_t tri = fbits(c.w);
    float4 ta = sc.tattr[2u * tri], tb = sc.tattr[2u * tri + 1u];
    uint32_t id = fbits(tb.w) >> 16;
    if (id != kTexNone) {
      float tu, tv;
      tri_texcoord(ta, tb, lo, 1.0f - hi, hi - lo, tu, tv);
      s.Le = s.Le * tex_sample(sc.tex[id], sc.texels, tu, tv);
    }
  }
  s.area = a.w;
  s.pdf_sel = b.w;
  s.idx = li;
  return s;
}

// Path-replay adjoint (no reference counterpart): the radiance still to come
// after a vertex, L_total - L_so_far, scales with that vertex's Lambertian
// albedo once per visit, and each emitted or light-sampled contribution scales
// with its light's radiance.  ps.adj_mode 0 is the primal pass (stores
// nothing extra); 1 replays the identical paths and scatters dLoss/dparam.
// Block-level accumulation of (key -> 3 floats) in LDS before the global
// float atomics.  The adjoint's derivatives land on few addresses (a handful
// of lights; the triangles a block's paths hit), whose HBM atomics would
// otherwise serialise.  A key that finds no slot in 8 probes goes to HBM.
constexpr uint32_t kAccEmpty = 0xFFFFFFFFu;
template <uint32_t N>
struct LdsAcc3 {
  uint32_t key[N];
  float val[N][3];
};

template <uint32_t N>
__device__ __forceinline__ void acc_init(LdsAcc3<N>& t) {
  for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
    t.key[i] = kAccEmpty;
    t.val[i][0] = t.val[i][1] = t.val[i][2] = 0.0f;
  }
}

template <uint32_t N>
__device__ __forceinline__ void acc_add(LdsAcc3<N>& t, float* dst, uint32_t key, V3 v) {
  uint32_t h = (key * 0x9E3779B1u) & (N - 1u);
  for (int probe = 0; probe < 8; ++probe) {
    uint32_t prev = atomicCAS(&t.key[h], kAccEmpty, key);
    if (prev == kAccEmpty || prev == key) {
      if (v.x != 0.0f) atomicAdd(&t.val[h][0], v.x);
      if (v.y != 0.0f) atomicAdd(&t.val[h][1], v.y);
      if (v.z != 0.0f) atomicAdd(&t.val[h][2], v.z);
      return;
    }
    h = (h + 1u) & (N - 1u);
  }
  if (v.x != 0.0f) unsafeAtomicAdd(dst + 3u * key + 0u, v.x);
  if (v.y != 0.0f) unsafeAtomicAdd(dst + 3u * key + 1u, v.y);
  if (v.z != 0.0f) unsafeAtomicAdd(dst + 3u * key + 2u, v.z);
}

template <uint32_t N>
__device__ __forceinline__ void acc_flush(LdsAcc3<N>& t, float* dst) {
  for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
    uint32_t k = t.key[i];
    if (k == kAccEmpty) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (t.val[i][c] != 0.0f) unsafeAtomicAdd(dst + 3u * k + (uint32_t)c, t.val[i][c]);
  }
}

// d L / d Le of light li for a contribution c = (...) * Le, per channel.
template <uint32_t N>
__device__ __forceinline__ void adj_add_le(LdsAcc3<N>& t, const PathSoA& ps, uint32_t li, V3 w, V3 c, V3 le) {
  V3 g{le.x > 0.0f ? w.x * c.x / le.x : 0.0f, le.y > 0.0f ? w.y * c.y / le.y : 0.0f,
       le.z > 0.0f ? w.z * c.z / le.z : 0.0f};
  acc_add(t, ps.grad_le, li, g);
}

__device__ __forceinline__ V3 adj_weight(const PathSoA& ps, uint32_t pid) {
  uint32_t pix = pid % ps.npix;
  return V3{ps.dldi[3u * pix], ps.dldi[3u * pix + 1u], ps.dldi[3u * pix + 2u]} * ps.inv_spp;
}

// Two queue appends with one global atomic per counter per work-group: the
// waves' counts are combined in LDS, so the hot counters see 8x fewer
// same-address atomics than with per-wave appends.
constexpr uint32_t kShadeBlock = 512;
constexpr uint32_t kShadeWaves = kShadeBlock / 64;

__device__ __forceinline__ void block_append2(uint32_t* ca, bool pa, uint32_t* cb, bool pb, uint32_t& sa,
                                              uint32_t& sb) {
  __shared__ uint32_t s_a[kShadeWaves + 1], s_b[kShadeWaves + 1];
  uint64_t ma = __ballot(pa), mb = __ballot(pb);
  uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  uint64_t lt = lane == 0 ? 0ull : ((~0ull) >> (64u - lane));
  uint32_t ba = (uint32_t)__popcll(ma & lt), bb = (uint32_t)__popcll(mb & lt);
  if (lane == 0) {
    s_a[wave] = (uint32_t)__popcll(ma);
    s_b[wave] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ta = 0, tb = 0;
    for (uint32_t w = 0; w < kShadeWaves; ++w) {
      uint32_t x = s_a[w], y = s_b[w];
      s_a[w] = ta;
      s_b[w] = tb;
      ta += x;
      tb += y;
    }
    s_a[kShadeWaves] = ta ? atomicAdd(ca, ta) : 0u;
    s_b[kShadeWaves] = tb ? atomicAdd(cb, tb) : 0u;
  }
  __syncthreads();
  sa = s_a[kShadeWaves] + s_a[wave] + ba;
  sb = s_b[kShadeWaves] + s_b[wave] + bb;
}

// kRis instantiates the RIS light selection; without it the kernel needs
// fewer registers and runs at a higher occupancy.
// SD-tree and hash-grid shading are held to 80 VGPRs (6 waves per SIMD, as
// plain shading; the hash grid's 82 gave 5 and took 8% longer).
// kTex instantiates texture and vertex-normal reads (scenes with attributes
// only: the code costs registers everywhere else).
template <int kGuide, bool kAdj, bool kRis, bool kTex>
constexpr unsigned shade_min_waves() {
  return ((kGuide == 1 || kGuide == 2) && !kAdj && !kRis && !kTex) ? 6u : 1u;
}

template <int kGuide, bool kAdj = false, bool kRis = false, bool kTex = false>
__global__ void __launch_bounds__(kShadeBlock)
__attribute__((amdgpu_waves_per_eu(shade_min_waves<kGuide, kAdj, kRis, kTex>())))
k_shade(RenderParams rp, SceneView sc, PathSoA ps, SdTreeView sd, HashGridView hg, MortonView mq, uint32_t depth) {
  __shared__ LdsAcc3<kAdj ? 1024u : 1u> acc_alb;
  __shared__ LdsAcc3<kAdj ? 256u : 1u> acc_le;
  const uint32_t b0 = blockIdx.x * kShadeBlock, i = b0 + threadIdx.x;
  const QueueRegion q = queue_region(ps, ps.count, depth, b0);
  if (b0 - q.base >= q.live) return;  // uniform over the work-group
  if (kAdj) {
    acc_init(acc_alb);
    acc_init(acc_le);
    __syncthreads();
  }
  bool active = i - q.base < q.live;
  bool cont = false, shadow = false;
  uint32_t pid = 0;
  V3 p{0, 0, 0}, nd{0, 0, 0}, sdir{0, 0, 0}, scon{0, 0, 0}, Tn{0, 0, 0};
  float wo_pdf = 0.0f, sdist = 0.0f;
  uint32_t slight = 0;
  uint32_t meta_next = 0;
  Pcg32 rng;
  const bool record = (kGuide != 0) && !(rp.flags & kFlagFinal);
  if (active) {
    pid = ps.queue[depth & 1][i];
    float4 ro = ps.ray_o[depth & 1][i], rd = ps.ray_d[depth & 1][i], T4 = ps.T[depth & 1][i], hit = ps.hit[i];
    uint2 rs = ps.rng[depth & 1][i];
    uint32_t npix = rp.width * rp.height;
    rng.state = ((uint64_t)rs.y << 32) | rs.x;
    rng.inc = pcg_stream_inc(pid % npix);
    uint32_t meta = fbits(rd.w);
    bool prev_delta = (meta & kMetaDelta) != 0;
    float prev_pdf = ro.w;
    V3 o{ro.x, ro.y, ro.z}, d{rd.x, rd.y, rd.z}, T{T4.x, T4.y, T4.z};
    V3 c{0.0f, 0.0f, 0.0f};
    float cos_out = 0.0f;
    int32_t tri = (int32_t)fbits(hit.w);
    const bool nee = (rp.flags & kFlagNee) && sc.n_lights > 0;
    const bool ris = kRis && nee && (rp.flags & kFlagRis);
    const bool mis = nee && !ris && (rp.flags & kFlagMis);
    uint32_t vcell = kInvalidCell;
    float vcu = 0.0f, vcv = 0.0f;
    uint32_t vflags = 0;
    const bool adj = kAdj && ps.adj_mode != 0u;
    V3 adj_w{0.0f, 0.0f, 0.0f};
    if (adj) adj_w = adj_weight(ps, pid);
    if (tri < 0) {
      c = T * V3{rp.bg[0], rp.bg[1], rp.bg[2]};
    } else {
      const float4* tq = reinterpret_cast<const float4*>(sc.tris + tri);
      float4 e1q = tq[1], e2q = tq[2];
      V3 ng = normalize(cross(load3(e1q), load3(e2q)));
      float dn = dot(d, ng);
      bool front = dn < 0.0f;
      V3 n = front ? ng : -ng;
      p = o + d * hit.x;
      V3 wo = -d;
      uint32_t info = sc.info[tri];
      if (tri_is_light(info)) {
        // Emission (SimpleTrace.cu:410-413) with MIS against light sampling.
        if (front) {
          uint32_t li = tri_light(info);
          const float4* L = sc.lights + 4u * li;
          float4 a = L[0], b = L[1], e = L[3];
          float w = 1.0f;
          if (depth > 0 && !prev_delta && nee) {
            if (mis) {
              float pl = b.w / a.w * hit.x * hit.x / fabsf(dn);
              w = power_heuristic(prev_pdf, pl);
            } else {
              w = 0.0f;
            }
          }
          V3 le{e.x, e.y, e.z};
          if (kTex && sc.tattr) {
            // SampleEmission at the hit's texture coordinate
            float4 ta = sc.tattr[2u * tri], tb = sc.tattr[2u * tri + 1u];
            float tu, tv;
            tri_texcoord(ta, tb, 1.0f - hit.y - hit.z, hit.y, hit.z, tu, tv);
            le = le * tex_factor(sc, fbits(tb.w) & 0xFFFFu, tu, tv);
          }
          c = T * le * w;
          if (adj) adj_add_le(acc_le, ps, li, adj_w, c, V3{e.x, e.y, e.z});
        }
      } else if (depth + 1u < rp.max_depth) {
        float4 sh = sc.shade[tri];
        uint32_t type = tri_type(info);
        V3 albedo{sh.x, sh.y, sh.z};
        float param = sh.w;
        if (adj && type == kMatLambert) {
          // This vertex's NEE and everything after it carry one factor of albedo.
          float4 lt = ps.adjL[pid], lc = ps.L[pid];
          V3 rest{lt.x - lc.x, lt.y - lc.y, lt.z - lc.z};
          acc_add(acc_alb, ps.grad_alb, (uint32_t)tri,
                  V3{albedo.x > 0.0f ? adj_w.x * rest.x / albedo.x : 0.0f,
                     albedo.y > 0.0f ? adj_w.y * rest.y / albedo.y : 0.0f,
                     albedo.z > 0.0f ? adj_w.z * rest.z / albedo.z : 0.0f});
        }
        V3 spec{0.0f, 0.0f, 0.0f};
        if (type == kMatPhong && sc.spec) {
          float4 s4 = sc.spec[tri];
          spec = V3{s4.x, s4.y, s4.z};
          param = s4.w;
        }
        float bu = hit.y, bv = hit.z;
        if (kTex && sc.tattr) {
          // SampleDiffuse / SampleSpecular at the hit's texture coordinate
          // (SimpleKernel.h:251-272); glass keeps its specular colour in albedo
          float4 ta = sc.tattr[2u * tri], tb = sc.tattr[2u * tri + 1u];
          float tu, tv;
          tri_texcoord(ta, tb, 1.0f - bu - bv, bu, bv, tu, tv);
          uint32_t ids = fbits(tb.z);
          albedo = albedo * tex_factor(sc, type == kMatGlass ? ids >> 16 : ids & 0xFFFFu, tu, tv);
          if (type == kMatPhong) spec = spec * tex_factor(sc, ids >> 16, tu, tv);
        }
        if (type == kMatGlass) {
          // Dielectric of SimpleTrace.cu:347-407 (Schlick Fresnel), a delta
          // event; it shades with the interpolated vertex normal when the
          // scene has one (GetTriangleVNormal, SimpleKernel.h:238-250)
          if (kTex && sc.vnorm) {
            const float4* vq = sc.vnorm + 3u * tri;
            float4 q0 = vq[0], q1 = vq[1], q2 = vq[2];
            float b0 = 1.0f - bu - bv;
            ng = normalize(V3{q0.x, q0.y, q0.z} * b0 + V3{q0.w, q1.x, q1.y} * bu + V3{q1.z, q1.w, q2.x} * bv);
          }
          float ior = param;
          float cos_i = dot(ng, d);
          V3 rn;
          float eta;
          if (cos_i < 0.0f) {
            rn = ng;
            eta = 1.0f / ior;
            cos_i = -cos_i;
          } else {
            rn = -ng;
            eta = ior;
          }
          float sin_o2 = (1.0f - cos_i * cos_i) * eta * eta;
          float f0 = (1.0f - eta) / (1.0f + eta);
          f0 = f0 * f0;
          float m = 1.0f - cos_i;
          float fres = f0 + (1.0f - f0) * (m * m * m * m * m);
          V3 refl = normalize(d - rn * (2.0f * dot(d, rn)));
          if (rnd(rng) < fres || sin_o2 > 1.0f) {
            p = p + rn * 1e-4f;
            nd = refl;
            Tn = T * albedo;
          } else {
            p = p - rn * 1e-4f;
            float sin_i2 = fmaxf(1.0f - cos_i * cos_i, 0.0f);
            float cos_o = sqrtf(1.0f - sin_o2);
            if (sin_i2 > 0.0f) {
              V3 k = (d + rn * cos_i) / sqrtf(sin_i2);
              nd = normalize(k * sqrtf(sin_o2) - rn * cos_o);
            } else {
              nd = d;
            }
            Tn = T;
          }
          wo_pdf = fabsf(dot(nd, rn));
          cos_out = wo_pdf;
          meta_next = (depth + 1u) | kMetaDelta;
          vflags = kVtxDelta;
          cont = max3(Tn) > 0.0f;
        } else {
          // vcell: where this vertex records; gcell: the guide it samples from
          uint32_t gcell = kInvalidCell;
          if (kGuide == 1) {
            vcell = stree_lookup(sd, p);
            gcell = vcell;
          } else if (kGuide == 2) {
            // SimpleGuide2.cu:319-323, 420-428: the guide is the previous
            // table's slot (FindFromPrv); the vertex records there, or, in a
            // Locate pass, into the slot a cell missing from the previous
            // table takes in the current one (FindFromCur)
            uint32_t ix, iy, iz, bucket, csum;
            grid_cell(hg, p, ix, iy, iz);
            grid_hash(hg, ix, iy, iz, bucket, csum);
            gcell = grid_find_prv(hg, bucket, csum);
            vcell = gcell;
            if ((rp.flags & kFlagLocate) && gcell == kInvalidCell) vcell = grid_find_cur(hg, bucket, csum);
          } else if (kGuide == 3 && record && sd.stree && records_path(ps, pid)) {
            // the learned guide's training vertices also record into an
            // SD-tree (its fallback), which does not guide this render
            vcell = stree_lookup(sd, p);
          }
          // learned guide: the mixture the inference pass wrote for this queue slot
          const float4* mx = nullptr;
          float csel = 0.0f;
          if (kGuide == 3 && (rp.flags & kFlagGuideSample) && depth < ps.mlp_guide_depth) {
            if (ps.mcache) {
              uint32_t e = mlp_cache_entry(ps.mcache_lo, ps.mcache_inv, ps.mcache_res, p, n);
              mx = ps.mcache + (size_t)e * kMlpMixF4;
              csel = ps.mcache_c[e];
            } else {
              mx = ps.mix + (size_t)i * kMlpMixF4;
              csel = ps.mixc[i];
            }
          }
          const bool guided = kGuide == 3 ? mx != nullptr
                                          : (kGuide != 0) && (rp.flags & kFlagGuideSample) && gcell != kInvalidCell;
          const float alpha = kGuide == 3 ? csel : rp.fraction;
          // ---- next-event estimation (SimpleTrace.cu:302-339); nee_state 1:
          // a light sample without MIS, 2: with MIS, completed after the
          // direction sampling
          uint32_t nee_state = 0u;
          float nee_pl = 0.0f, nee_bp = 0.0f, nee_cos = 0.0f;
          V3 nee_f{0.0f, 0.0f, 0.0f};
          if (nee) {
            if (ris) {
              float wsum = 0.0f, fa_y = 0.0f, dist_y = 0.0f;
              uint32_t m = 0;
              V3 f_y{0, 0, 0}, dir_y{0, 0, 0};
              for (uint32_t k = 0; k < rp.num_candidates; ++k) {
                LightSample ls = sample_light<kTex>(sc, rng);
                V3 dl = ls.p - p;
                float dist = length(dl);
                V3 dir = dl / dist;
                float ndl = dot(dir, n);
                float lndl = -dot(dir, ls.n);
                BsdfEval be = bsdf_eval(type, albedo, param, spec, n, wo, dir);
                float g = fmaxf(ndl, 0.0f) * fmaxf(lndl, 0.0f) / (dist * dist);
                V3 f = be.f * ls.Le * g;
                float fa = avg3(f);
                float wi = fa * (ls.area / ls.pdf_sel);
                wsum += wi;
                ++m;
                float ur = rnd(rng);
                if ((wi / wsum) >= ur) {
                  f_y = f;
                  fa_y = fa;
                  dir_y = dir;
                  dist_y = dist;
                  slight = ls.idx;
                }
              }
              if (wsum > 0.0f && fa_y > 0.0f) {
                float W = wsum / (fa_y * (float)m);
                scon = T * f_y * W;
                sdir = dir_y;
                sdist = dist_y - 1e-4f;
                shadow = max3(scon) > 0.0f;
              }
            } else {
              LightSample ls = sample_light<kTex>(sc, rng);
              V3 dl = ls.p - p;
              float dist2 = dot(dl, dl);
              float dist = sqrtf(dist2);
              V3 dir = dl / dist;
              float cos_l = -dot(dir, ls.n);
              float cos_s = dot(dir, n);
              if (cos_l > 0.0f && cos_s > 0.0f) {
                // the MIS weight waits for the guide's density of this
                // direction, which is evaluated together with the sampling
                // below
                BsdfEval be = bsdf_eval(type, albedo, param, spec, n, wo, dir);
                nee_pl = ls.pdf_sel / ls.area * dist2 / cos_l;
                nee_bp = be.pdf;
                nee_cos = cos_s;
                nee_f = T * be.f * ls.Le;
                nee_state = mis ? 2u : 1u;
                sdir = dir;
                sdist = dist - 1e-4f;
                slight = ls.idx;
              }
            }
          }
          // ---- direction sampling: BSDF or guide mixture (SimpleGuide.cu:359-377)
          V3 dir;
          // the guide's density of the continuation direction and of the
          // light-sample direction (needed when NEE weights by MIS)
          float gp = 0.0f, gp_nee = 0.0f;
          const bool need_nee = guided && nee_state == 2u;
          if (kGuide == 1 || kGuide == 2) {
            const bool gsample = guided && rnd(rng) < alpha;
            if (!gsample) dir = bsdf_sample(type, albedo, param, spec, n, wo, rng);
            if (guided) {
              // a direction drawn from the tree, or the BSDF direction's
              // density, in the same descent loop as the light direction's
              // density (SD-tree), or the same load round (hash grid:
              // MortonQuadTreeWrapperT::SampleAndPdf, OPX7UtilsMorton.h:486-500,
              // gives the density of the sampled bin)
              float ncu = 0.0f, ncv = 0.0f;
              if (need_nee) dir_to_canonical(sdir, ncu, ncv);
              if (!gsample) dir_to_canonical(dir, vcu, vcv);
              float gu = vcu, gv = vcv;
              if (kGuide == 1) {
                dtree_pair(sd, gcell, gsample, rng, gu, gv, gp, need_nee, ncu, ncv, gp_nee);
              } else {
                mq_pair(mq, gcell, gsample, rng, gu, gv, gp, need_nee, ncu, ncv, gp_nee);
              }
              if (gsample) {
                dir = normalize(canonical_to_dir(gu, gv));
                // the recorded coordinates are those of the normalised
                // direction (only a recording pass reads them)
                if (record) dir_to_canonical(dir, vcu, vcv);
              }
            }
          } else if (guided) {
            dir = rnd(rng) < alpha ? mix_sample(mx, rng) : bsdf_sample(type, albedo, param, spec, n, wo, rng);
            gp = mix_pdf(mx, dir);
            if (need_nee) gp_nee = mix_pdf(mx, sdir);
            if (vcell != kInvalidCell) dir_to_canonical(dir, vcu, vcv);
          } else {
            dir = bsdf_sample(type, albedo, param, spec, n, wo, rng);
          }
          if (nee_state != 0u) {
            float w = 1.0f;
            if (nee_state == 2u) {
              float bp = nee_bp;
              if (guided) bp = (1.0f - alpha) * bp + alpha * gp_nee;
              w = power_heuristic(nee_pl, bp);
            }
            scon = nee_f * (nee_cos * w / nee_pl);
            shadow = max3(scon) > 0.0f;
          }
          BsdfEval be = bsdf_eval(type, albedo, param, spec, n, wo, dir);
          wo_pdf = be.pdf;
          if (guided) {
            wo_pdf = (1.0f - alpha) * be.pdf + alpha * fmaxf(gp, 0.0f);
          } else if (record && vcell != kInvalidCell) {
            dir_to_canonical(dir, vcu, vcv);
          }
          float cosd = dot(dir, n);
          cos_out = cosd;
          Tn = (wo_pdf > 0.0f && cosd > 0.0f) ? T * be.f * (cosd / wo_pdf) : V3{0.0f, 0.0f, 0.0f};
          if ((rp.flags & kFlagRR) && depth >= rp.rr_depth) {
            float q = fminf(max3(Tn), 0.95f);
            if (rnd(rng) >= q) {
              Tn = V3{0.0f, 0.0f, 0.0f};
            } else {
              Tn = Tn / q;
            }
          }
          bool bsdf_ok = finite3(be.f) && be.f.x >= 0.0f && be.f.y >= 0.0f && be.f.z >= 0.0f;
          vflags = bsdf_ok ? 0u : kVtxBadBsdf;
          if (kGuide == 3 && record && records_path(ps, pid)) {
            // training vertex: network inputs, sampled direction, BSDF density
            // and the BSDF weight f |cos| / p_sample before Russian roulette
            size_t mi = (size_t)depth * ps.mlp_rec + pid / ps.mlp_stride, st = (size_t)rp.max_depth * ps.mlp_rec;
            float rough = type == kMatGGX ? param : 0.0f;
            V3 bw = (wo_pdf > 0.0f) ? be.f * (fabsf(cosd) / wo_pdf) : V3{0.0f, 0.0f, 0.0f};
            uint32_t ok = (wo_pdf > 0.0f && bsdf_ok) ? kVtxValid : 0u;
            ps.mv[mi] = make_float4(p.x, p.y, p.z, rough);
            ps.mv[st + mi] = make_float4(n.x, n.y, n.z, be.pdf);
            ps.mv[2 * st + mi] = make_float4(wo.x, wo.y, wo.z, bitsf(ok));
            ps.mv[3 * st + mi] = make_float4(dir.x, dir.y, dir.z, 0.0f);
            ps.mv[4 * st + mi] = make_float4(bw.x, bw.y, bw.z, 0.0f);
          }
          nd = dir;
          meta_next = depth + 1u;
          cont = max3(Tn) > 0.0f && isfinite(max3(Tn));
        }
      }
    }
    // Path radiance (read and written only when this hit adds to it) and, for
    // guiding, the per-hit contribution.
    if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
      float4 L4 = ps.L[pid];
      ps.L[pid] = make_float4(L4.x + c.x, L4.y + c.y, L4.z + c.z, 0.0f);
    }
    if (record && records_path(ps, pid)) {
      ps.nhit[pid] = depth + 1u;
      size_t vi = (size_t)depth * ps.n + pid;
      ps.vB[vi] = make_float4(c.x, c.y, c.z, cos_out);
      // a path's last hit is never committed (only its contribution, vB, is
      // read): its throughput and guide records are not written
      if (cont) ps.vA[vi] = make_float4(Tn.x, Tn.y, Tn.z, wo_pdf);
      if (cont && (kGuide != 3 || sd.stree)) {
        bool has_vertex = cont && vcell != kInvalidCell;
        if ((kGuide == 1 || kGuide == 3) && sd.filter && has_vertex) {
          // stochastic box filters: the record moves to a random point of the
          // reference's boxes (own stream); u < 0 marks a dropped flux
          uint32_t npix = rp.width * rp.height;
          Pcg32 fr = filter_rng(rp.seed, pid % npix, rp.sample_offset + pid / npix, depth);
          float f0 = rnd(fr), f1 = rnd(fr), f2 = rnd(fr), f3 = rnd(fr), f4 = rnd(fr);
          if (sd.filter & kFilterSpatialBox) vcell = stree_box_jitter(sd, p, vcell, f0, f1, f2);
          has_vertex = vcell != kInvalidCell;
          if (has_vertex && (sd.filter & kFilterDirectionalBox) && !dtree_box_jitter(sd, vcell, vcu, vcv, f3, f4)) {
            vcu = -1.0f;
          }
        }
        // integer store: the all-ones sentinel is a NaN pattern as a float
        reinterpret_cast<uint4*>(ps.vC)[vi] = make_uint4(fbits(vcu), fbits(vcv), has_vertex ? vcell : kInvalidCell,
                                                         vflags | (has_vertex ? kVtxValid : 0u));
      }
      if (kGuide == 3 && (!cont || (vflags & kVtxDelta))) {
        // no training vertex here (light, miss, delta or absorbed path)
        size_t st = (size_t)rp.max_depth * ps.mlp_rec;
        ps.mv[2 * st + (size_t)depth * ps.mlp_rec + pid / ps.mlp_stride] = make_float4(0.0f, 0.0f, 0.0f, bitsf(0u));
      }
    }
  }
  uint32_t slot, sslot;
  block_append2(&ps.count[(depth + 1u) * kQueueRegions + q.index], cont, &ps.scount[depth * kQueueRegions + q.index],
                shadow, slot, sslot);
  slot += q.base;
  sslot += q.base;
  if (kAdj) {
    __syncthreads();
    acc_flush(acc_alb, ps.grad_alb);
    acc_flush(acc_le, ps.grad_le);
  }
  if (cont) {
    // the path's state moves to its slot in the next depth's queue
    const uint32_t nb = (depth + 1u) & 1u;
    ps.queue[nb][slot] = pid;
    ps.ray_o[nb][slot] = make_float4(p.x, p.y, p.z, wo_pdf);
    ps.ray_d[nb][slot] = make_float4(nd.x, nd.y, nd.z, bitsf(meta_next));
    ps.T[nb][slot] = make_float4(Tn.x, Tn.y, Tn.z, 0.0f);
    ps.rng[nb][slot] = make_uint2((uint32_t)rng.state, (uint32_t)(rng.state >> 32));
  }
  if (shadow) {
    float4* e = ps.shadow + 3u * (size_t)sslot;
    e[0] = make_float4(p.x, p.y, p.z, sdist);
    e[1] = make_float4(sdir.x, sdir.y, sdir.z, bitsf(pid));
    e[2] = make_float4(scon.x, scon.y, scon.z, bitsf(slight));
  }
}

// Any-hit traversal held to 8 waves per SIMD like k_extend.
template <bool kCount, bool kAdj = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kAdj ? 1 : 8)))
k_shadow(SceneView sc, PathSoA ps, uint32_t depth, uint32_t record) {
  __shared__ uint2 lds[kShortStack][kBlock];
  __shared__ LdsAcc3<kAdj ? 256u : 1u> acc_le;
  const uint32_t b0 = blockIdx.x * kBlock, i = b0 + threadIdx.x;
  const QueueRegion q = queue_region(ps, ps.scount, depth, b0);
  if (b0 - q.base >= q.live) return;  // uniform over the work-group
  if (kAdj) {
    acc_init(acc_le);
    __syncthreads();
  }
  uint32_t cnt[2] = {0u, 0u};
  bool occ = true;
  uint32_t pid = 0;
  const float4* e = ps.shadow + 3u * (size_t)i;
  if (i - q.base < q.live) {
    float4 a = e[0], b = e[1];
    pid = fbits(b.w);
    StackRef st = make_stack(lds, ps.deep, ps.n, i, sc.max_stack);
    float t, u, v;
    int32_t tri;
    occ = traverse<true, kCount>(sc, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, kRayTmin, a.w, st, ps.err, t, u, v, tri,
                                 cnt);
  }
  if (kCount) wave_count(ps.work + 2, cnt[0], cnt[1]);
  if (!occ) {
    float4 c = e[2];  // contribution, loaded after the traversal to keep it out of its registers
    float4 L4 = ps.L[pid];
    ps.L[pid] = make_float4(L4.x + c.x, L4.y + c.y, L4.z + c.z, 0.0f);
    if (kAdj && ps.adj_mode != 0u) {
      uint32_t li = fbits(c.w);
      float4 le = sc.lights[4u * li + 3u];
      adj_add_le(acc_le, ps, li, adj_weight(ps, pid), V3{c.x, c.y, c.z}, V3{le.x, le.y, le.z});
    }
    if (record && records_path(ps, pid)) {
      size_t vi = (size_t)depth * ps.n + pid;
      float4 vb = ps.vB[vi];
      ps.vB[vi] = make_float4(vb.x + c.x, vb.y + c.y, vb.z + c.z, vb.w);
    }
  }
  if (kAdj) {
    __syncthreads();
    acc_flush(acc_le, ps.grad_le);
  }
}

// Depth 0: region k holds the camera paths of its slots.
__global__ void k_set_count(uint32_t* c, uint32_t n, uint32_t qreg, uint32_t nreg) {
  uint32_t k = threadIdx.x;
  if (k < nreg) c[k] = min(qreg, n - k * qreg);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Sums the launch's samples per pixel in sample order and adds to the image.
__global__ void k_accum(PathSoA ps, uint32_t npix, uint32_t spp, float* accum) {
  uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= npix) return;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f, m1 = 0.0f, m2 = 0.0f;
  for (uint32_t s = 0; s < spp; ++s) {
    float4 l = ps.L[(size_t)s * npix + pix];
    sx += l.x;
    sy += l.y;
    sz += l.z;
    float y = (l.x + l.y + l.z) / 3.0f;
    m1 += y;
    m2 += y * y;
  }
  accum[3 * pix + 0] += sx;
  accum[3 * pix + 1] += sy;
  accum[3 * pix + 2] += sz;
  if (ps.moments) {
    ps.moments[2 * pix] += m1;
    ps.moments[2 * pix + 1] += m2;
  }
}

// Display frame of the accumulated image (SimpleTrace.cu:91-101): x = sum /
// spp, x / (1 + x), the reference's gamma approximation (Math.h:621-627:
// v <= 0.00174 ? 32 v : v^(1/2.2)), then min(int(255.99 g), 255); alpha 255.
__global__ void k_tonemap(const float* accum, uint32_t npix, float inv_spp, uchar4* frame) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  unsigned char c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float x = accum[3 * i + k] * inv_spp;
    x = x / (1.0f + x);
    float g = (x <= 0.00174f) ? (32.0f * x) : powf(x, 1.0f / 2.2f);
    int v = (int)(g * 255.99f);
    c[k] = (unsigned char)(v < 255 ? (v > 0 ? v : 0) : 255);
  }
  frame[i] = make_uchar4(c[0], c[1], c[2], 255);
}

// ------------------------------------------------------------------ debug frames
// __raygen__debug / __closesthit__debug / __miss__debug (SimpleTrace.cu:104-213,
// 631-673, 714-734): one camera ray per pixel, jittered by the first two numbers
// of the stream of sample rp.sample_offset (as k_camera), closest hit in
// [1e-4, 1e20], and per pixel one quantity (DebugFrameType, SimpleKernel.h:
// 160-171; 8 = world position has no reference counterpart).  val: the raw
// quantity and the hit primitive (original index bits, ~0 on a miss); frame:
// the reference's display mapping (normal (n+1)/2, depth t/(1+t), shininess
// s/(1+s), index r/(1+r)), linear_to_gamma, x 255.99, clamped.  A miss gives
// normal (1,0,0), depth 1e20 and zeros.
__device__ __forceinline__ unsigned char debug_byte(float x) {
  float g = (x <= 0.00174f) ? (32.0f * x) : powf(x, 1.0f / 2.2f);
  int v = (int)(g * 255.99f);
  return (unsigned char)(v < 255 ? (v > 0 ? v : 0) : 255);
}

__global__ void __launch_bounds__(kBlock) k_debug(RenderParams rp, Camera cam, SceneView sc, HashGridView hg,
                                                  MortonView mq, uint32_t type, uint32_t use_grid, float4* val,
                                                  uchar4* frame, uint32_t* deep, uint32_t* err) {
  __shared__ uint2 lds[kShortStack][kBlock];
  uint32_t npix = rp.width * rp.height;
  uint32_t pix = blockIdx.x * kBlock + threadIdx.x;
  if (pix >= npix) return;
  uint32_t x = pix % rp.width, y = pix / rp.width;
  Pcg32 rng = path_rng(rp.seed, pix, rp.sample_offset);
  float jx = rnd(rng), jy = rnd(rng);
  float dx = 2.0f * ((float)x + jx) / (float)rp.width - 1.0f;
  float dy = 2.0f * ((float)y + jy) / (float)rp.height - 1.0f;
  V3 U{cam.u[0], cam.u[1], cam.u[2]}, Vv{cam.v[0], cam.v[1], cam.v[2]}, W{cam.w[0], cam.w[1], cam.w[2]};
  V3 d = normalize(dx * U + dy * Vv + W);
  V3 o{cam.eye[0], cam.eye[1], cam.eye[2]};
  StackRef st = make_stack(lds, deep, npix, pix, sc.max_stack);
  float t, u, v;
  int32_t tri;
  traverse<false>(sc, o, d, 1e-4f, 1e20f, st, err, t, u, v, tri);
  V3 q{0.0f, 0.0f, 0.0f};
  uint32_t prim = 0xFFFFFFFFu;
  if (tri < 0) {
    if (type == 0) q = V3{1.0f, 0.0f, 0.0f};
    if (type == 1) q = V3{1e20f, 1e20f, 1e20f};
  } else {
    const float4* tq = reinterpret_cast<const float4*>(sc.tris + tri);
    prim = fbits(tq[0].w);
    V3 p = o + d * t;
    float4 sh = sc.shade[tri];
    uint32_t info = sc.info[tri];
    uint32_t mt = tri_type(info);
    // texture factors of the diffuse, specular and emission slots
    float tu = 0.0f, tv = 0.0f;
    uint32_t ids = 0xFFFFFFFFu, ide = 0xFFFFFFFFu;
    if (sc.tattr && type >= 2u && type <= 4u) {
      float4 ta = sc.tattr[2u * tri], tb = sc.tattr[2u * tri + 1u];
      tri_texcoord(ta, tb, 1.0f - u - v, u, v, tu, tv);
      ids = fbits(tb.z);
      ide = fbits(tb.w);
    }
    switch (type) {
      case 0:
        q = normalize(cross(load3(tq[1]), load3(tq[2])));
        break;
      case 1:
        q = V3{t, t, t};
        break;
      case 2:
        if (mt != kMatGlass && !tri_is_light(info)) q = V3{sh.x, sh.y, sh.z} * tex_factor(sc, ids & 0xFFFFu, tu, tv);
        break;
      case 3:
        if (mt == kMatGlass) {
          q = V3{sh.x, sh.y, sh.z} * tex_factor(sc, ids >> 16, tu, tv);
        } else if (mt == kMatPhong && sc.spec) {
          float4 s4 = sc.spec[tri];
          q = V3{s4.x, s4.y, s4.z} * tex_factor(sc, ids >> 16, tu, tv);
        }
        break;
      case 4:
        if (tri_is_light(info)) {
          float4 e = sc.lights[4u * tri_light(info) + 3u];
          q = V3{e.x, e.y, e.z} * tex_factor(sc, ide & 0xFFFFu, tu, tv);
        }
        break;
      case 5:
        if (mt == kMatPhong && sc.spec) q = V3{sc.spec[tri].w, sc.spec[tri].w, sc.spec[tri].w};
        break;
      case 6:
        if (mt == kMatGlass) q = V3{sh.w, sh.w, sh.w};
        break;
      case 7:
        if (use_grid) {
          // root weight of the building direction tree of the hit cell
          uint32_t ix, iy, iz, bucket, csum;
          grid_cell(hg, p, ix, iy, iz);
          grid_hash(hg, ix, iy, iz, bucket, csum);
          uint32_t cell = grid_find_prv(hg, bucket, csum);
          if (cell != kInvalidCell) {
            float w = mq.build[(size_t)mq.per_cell * cell];
            q = V3{w, w, w};
          }
        }
        break;
      default:
        q = p;
        break;
    }
  }
  // integer store: the miss marker (all ones) is a NaN pattern as a float
  if (val) reinterpret_cast<uint4*>(val)[pix] = make_uint4(fbits(q.x), fbits(q.y), fbits(q.z), prim);
  if (frame) {
    V3 c = q;
    if (type == 0) c = (q + V3{1.0f, 1.0f, 1.0f}) * 0.5f;
    if (type == 1 || type == 5 || type == 6) c = V3{q.x / (1.0f + q.x), q.y / (1.0f + q.y), q.z / (1.0f + q.z)};
    frame[pix] = make_uchar4(debug_byte(c.x), debug_byte(c.y), debug_byte(c.z), 255);
  }
}

hipError_t launch_debug(const RenderParams& rp, const Camera& cam, const SceneView& sc, const HashGridView& hg,
                        const MortonView& mq, uint32_t type, bool use_grid, float4* val, uchar4* frame,
                        uint32_t* deep, uint32_t* err, hipStream_t s) {
  uint32_t npix = rp.width * rp.height;
  if (npix == 0) return hipSuccess;
  hipLaunchKernelGGL(k_debug, dim3((npix + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, rp, cam, sc, hg, mq, type,
                     use_grid ? 1u : 0u, val, frame, deep, err);
  return hipGetLastError();
}

hipError_t launch_tonemap(const float* accum, uint32_t npix, float inv_spp, uchar4* frame, hipStream_t s) {
  if (npix == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tonemap, dim3((npix + 255u) / 256u), dim3(256), 0, s, accum, npix, inv_spp, frame);
  return hipGetLastError();
}

// Per-sample relative variance of a moments render, per work-group of 256
// pixels: sum of s^2 / (m^2 + 0.01), s^2 the unbiased variance of the
// render's n samples of y, m the pixel's y in the accumulated image.
__global__ void k_relvar(const float* mom, const float* accum, uint32_t npix, float n, float inv_acc,
                         double* partial) {
  __shared__ double red[256];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < npix) {
    double s1 = mom[2 * i], s2 = mom[2 * i + 1];
    double var = fmax(0.0, (s2 - s1 * s1 / n) / (n - 1.0));
    double ref = ((double)accum[3 * i] + accum[3 * i + 1] + accum[3 * i + 2]) * inv_acc;
    v = var / (ref * ref + 1e-2);
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

hipError_t launch_relvar(const float* mom, const float* accum, uint32_t npix, uint32_t n, uint32_t acc_spp,
                         double* partial, hipStream_t s) {
  if (npix == 0) return hipSuccess;
  hipLaunchKernelGGL(k_relvar, dim3((npix + 255u) / 256u), dim3(256), 0, s, mom, accum, npix, (float)n,
                     1.0f / (3.0f * (float)(acc_spp ? acc_spp : 1u)), partial);
  return hipGetLastError();
}

// ------------------------------------------------------------------ guide records
// One kernel commits the guide vertices of a chunk (TraceVertexT::Commit,
// OPX7UtilsPathGuiding.h:830-866; MortonTraceVertexT::Commit,
// OPX7UtilsMorton.h:550-585).  A work-item takes one path at a time and walks
// its vertices from the last hit backwards, so the radiance that reached
// vertex d (the contributions of hits d+1 .. nh-1) is one running sum (the
// reference sums them forwards per vertex: same terms, other rounding).  Each
// committed vertex adds 1 to its DTree's statistical weight and its value to
// the leaf slot of its direction (hash grid: to the Morton leaf of its slot).
// Both go through LDS tables of the work-group, flushed with one global float
// atomic per distinct address, so hot DTrees and leaves near the quadtree
// roots do not serialise on HBM atomics; a key that finds no LDS slot in 16
// probes goes to HBM directly.
constexpr int kMaxCommitVtx = 16;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
// Work-group shape per guide: work-items, paths per work-item, LDS slots for
// leaf values and for DTree weights.  The SD-tree's leaves are many and
// spread (a DTree each per spatial leaf): 16 waves over 8192 paths share a
// 16k-slot table (144 KB of LDS, one work-group per CU), which misses far
// less often than four waves with 4k slots (measured: 9.0 vs 12.2 ms per
// 64-spp recording pass of the room).  The hash grid's Morton leaves fit the
// smaller shape.
template <int kKind>
struct CommitShape;
// probes: LDS slots tried before a record goes to global memory (the hash
// grid's small table is often full, where long probe runs only cost time:
// 4 probes instead of 16 measured -7% on its commit).
template <>
struct CommitShape<1> {
  static constexpr uint32_t block = 1024, paths = 8, leaf = 16384, cell = 2048;
  static constexpr int probes = 16;
};
template <>
struct CommitShape<2> {
  static constexpr uint32_t block = 256, paths = 32, leaf = 4096, cell = 1;
  static constexpr int probes = 4;
};

// Insert-or-add into an LDS open-addressing table; false when kProbes probes
// miss (the caller then adds to global memory directly).
template <int kProbes, class T>
__device__ __forceinline__ bool lds_table_add(uint32_t* keys, T* vals, uint32_t mask, uint32_t key, T v) {
  uint32_t h = (key * 0x9E3779B1u) & mask;
  for (int probe = 0; probe < kProbes; ++probe) {
    uint32_t prev = atomicCAS(&keys[h], kEmptyKey, key);
    if (prev == kEmptyKey || prev == key) {
      atomicAdd(&vals[h], v);
      return true;
    }
    h = (h + 1u) & mask;
  }
  return false;
}

// Flux slot of a record (DTreeNodeT::Record :113-125): the first levels of the
// descent from the record jump table (cell = top bits of u, v; the remaining
// coordinates are exact, as in dpdf_jump_slot), the rest node by node.
__device__ __forceinline__ uint32_t record_key(const SdTreeView& sd, uint32_t cell, float cu, float cv) {
  const float sn = (float)(1u << kPdfJumpBits);
  float gx = cu * sn, gy = cv * sn;
  float jx = fminf(fmaxf(floorf(gx), 0.0f), sn - 1.0f), jy = fminf(fmaxf(floorf(gy), 0.0f), sn - 1.0f);
  uint2 e = sd.build_jump[(size_t)cell * kPdfJumpCells + ((uint32_t)jy << kPdfJumpBits) + (uint32_t)jx];
  const uint32_t base = e.y;
  if (e.x & 0x80000000u) return base * 4u + (e.x & 0x7FFFFFFFu);
  cu = gx - jx;
  cv = gy - jy;
  uint32_t node = e.x;
  int idx = dtree_child_idx(cu, cv);
  for (;;) {
    uint32_t ch = dnode_child(sd.build_child[base + node], idx);
    if (ch == 0) break;
    node = ch;
    idx = dtree_child_idx(cu, cv);
  }
  return (base + node) * 4u + (uint32_t)idx;
}

// Leaf bin of a record in the Morton quadtree of hash cell `cell`.
__device__ __forceinline__ uint32_t morton_key(const MortonView& mq, uint32_t cell, float cu, float cv) {
  uint32_t nn = 1u << mq.level;
  uint32_t x = (uint32_t)(cu * (float)nn), y = (uint32_t)(cv * (float)nn);
  x = x < nn ? x : nn - 1u;
  y = y < nn ? y : nn - 1u;
  return mq.per_cell * cell + mq_level_offset(mq.level) + morton2(x, y);
}

// kind 1: SD-tree (DTreeT::RecordIrradiance :419-427 -> DTreeNodeT::Record
// :113-125: weight += 1, flux into the first leaf slot on the descent, for a
// positive finite value not dropped by the directional box filter, cu < 0);
// kind 2: Morton quadtree leaf (MortonQuadTreeT::Record :262-285, which tests
// no value).
template <int kKind>
__global__ void __launch_bounds__(CommitShape<kKind>::block) k_commit(PathSoA ps, uint32_t max_vtx, SdTreeView sd,
                                                                     MortonView mq) {
  constexpr uint32_t kCommitBlock = CommitShape<kKind>::block, kCommitPaths = CommitShape<kKind>::paths;
  constexpr uint32_t kLeafTab = CommitShape<kKind>::leaf, kCellTab = CommitShape<kKind>::cell;
  __shared__ uint32_t lkey[kLeafTab];
  __shared__ float lval[kLeafTab];
  // DTree weights: with at most 2 kCellTab DTrees a count per DTree (no
  // probing), else an open-addressing table of kCellTab keys and counts
  __shared__ uint32_t cwords[2 * kCellTab];
  uint32_t* const ckey = cwords;
  uint32_t* const cval = cwords + kCellTab;
  const bool cdense = kKind == 1 && sd.n_dtrees <= 2u * kCellTab;
  const uint32_t first = blockIdx.x * kCommitBlock * kCommitPaths;
  if (first >= ps.n) return;  // uniform over the work-group
  for (uint32_t i = threadIdx.x; i < kLeafTab; i += kCommitBlock) {
    lkey[i] = kEmptyKey;
    lval[i] = 0.0f;
  }
  if (kKind == 1) {
    for (uint32_t i = threadIdx.x; i < kCellTab; i += kCommitBlock) {
      ckey[i] = cdense ? 0u : kEmptyKey;
      cval[i] = 0u;
    }
  }
  __syncthreads();
  float* flat = kKind == 1 ? reinterpret_cast<float*>(sd.build_sums) : mq.build;
  for (uint32_t k = 0; k < kCommitPaths; ++k) {
    // consecutive work-items take consecutive paths, so the vertex loads of a
    // wave at one depth are contiguous
    const uint32_t pid = first + k * kCommitBlock + threadIdx.x;
    const bool in = pid < ps.n && records_path(ps, pid);
    uint32_t nh = in ? ps.nhit[pid] : 0u;
    nh = nh < (uint32_t)kMaxCommitVtx ? nh : (uint32_t)kMaxCommitVtx;
    uint32_t nv = nh > 0 ? nh - 1u : 0u;
    nv = nv < max_vtx ? nv : max_vtx;
    uint32_t nh_max = nh;
    for (int off = 32; off >= 1; off >>= 1) nh_max = max(nh_max, (uint32_t)__shfl_xor((int)nh_max, off));
    float rx = 0.0f, ry = 0.0f, rz = 0.0f;
    // vB record of hit d + 1 (its contribution); read once per hit, it then
    // gives the cosine of vertex d + 1 and the contribution at d
    float4 b = nh >= 2u ? ps.vB[(size_t)(nh - 1u) * ps.n + pid] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int d = (int)nh_max - 2; d >= 0; --d) {
      if ((uint32_t)d + 1u < nh) {
        rx += b.x;
        ry += b.y;
        rz += b.z;
        b = ps.vB[(size_t)d * ps.n + pid];
      }
      if ((uint32_t)d >= nv) continue;
      const size_t vi = (size_t)d * ps.n + pid;
      const float4 vc = ps.vC[vi];
      const uint32_t fl = fbits(vc.w), cell = fbits(vc.z);
      if (!(fl & kVtxValid) || (fl & kVtxDelta) || cell == kInvalidCell) continue;
      const float4 va = ps.vA[vi];
      const float wo_pdf = va.w;
      const bool fin = isfinite(rx) && rx >= 0.0f && isfinite(ry) && ry >= 0.0f && isfinite(rz) && rz >= 0.0f;
      if (!(wo_pdf > 0.0f && fin && !(fl & kVtxBadBsdf))) continue;
      float lx = va.x * wo_pdf > 1e-4f ? rx / va.x : 0.0f;
      float ly = va.y * wo_pdf > 1e-4f ? ry / va.y : 0.0f;
      float lz = va.z * wo_pdf > 1e-4f ? rz / va.z : 0.0f;
      const float cosv = fabsf(b.w);  // b: this vertex's record
      lx *= cosv;
      ly *= cosv;
      lz *= cosv;
      const float value = ((lx + ly + lz) / 3.0f) / wo_pdf;
      const float cu = vc.x, cv = vc.y;
      if (kKind == 1) {
        if (cdense) {
          atomicAdd(&cwords[cell], 1u);
        } else if (!lds_table_add<CommitShape<kKind>::probes, uint32_t>(ckey, cval, kCellTab - 1u, cell, 1u)) {
          unsafeAtomicAdd(sd.build_sw + cell, 1.0f);
        }
        if (!(isfinite(value) && value > 0.0f && cu >= 0.0f)) continue;
      }
      const uint32_t key = kKind == 1 ? record_key(sd, cell, cu, cv) : morton_key(mq, cell, cu, cv);
      if (!lds_table_add<CommitShape<kKind>::probes, float>(lkey, lval, kLeafTab - 1u, key, value)) unsafeAtomicAdd(flat + key, value);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kLeafTab; i += kCommitBlock) {
    if (lkey[i] != kEmptyKey) unsafeAtomicAdd(flat + lkey[i], lval[i]);
  }
  if (kKind == 1) {
    if (cdense) {
      for (uint32_t i = threadIdx.x; i < sd.n_dtrees; i += kCommitBlock) {
        if (cwords[i] != 0u) unsafeAtomicAdd(sd.build_sw + i, (float)cwords[i]);
      }
    } else {
      for (uint32_t i = threadIdx.x; i < kCellTab; i += kCommitBlock) {
        if (ckey[i] != kEmptyKey) unsafeAtomicAdd(sd.build_sw + ckey[i], (float)cval[i]);
      }
    }
  }
}

// ------------------------------------------------------------------ learned guide
// Inference pass of one bounce: for every live path of the queue, the network
// maps (hit point, direction towards the previous vertex, front-facing normal,
// roughness) to the vMF mixture and the selection probability, written per
// queue slot for k_shade.  Persistent work-groups of 4 waves x 16 paths keep
// the bf16 weights in LDS.
constexpr uint32_t kInferBlock = 256;
constexpr uint32_t kInferGrid = 2048;

__global__ void __launch_bounds__(kInferBlock) k_mlp_infer(SceneView sc, PathSoA ps, MlpView m, uint32_t depth) {
  __shared__ uint16_t lw[kLdsWeights];
  __shared__ float hs[4][16 * kHeadPitch];
  const uint32_t nq = ps.nreg * ps.qreg;
  mlp_load_weights(lw, m.w);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15;
  float* row = hs[wave] + r * kHeadPitch;
  for (uint32_t tile = blockIdx.x; tile * 64u < nq; tile += gridDim.x) {
    const QueueRegion q = queue_region(ps, ps.count, depth, tile * 64u);
    if (tile * 64u - q.base >= q.live) continue;  // uniform over the work-group
    uint32_t i = tile * 64u + 16u * (uint32_t)wave + (uint32_t)r;
    bool valid = i - q.base < q.live;
    MlpInput in{V3{0, 0, 0}, V3{0, 0, 1}, V3{0, 0, 1}, 0.0f};
    if (valid) {
      float4 hit = ps.hit[i];
      int32_t tri = (int32_t)fbits(hit.w);
      if (tri >= 0) {
        float4 ro = ps.ray_o[depth & 1][i], rd = ps.ray_d[depth & 1][i];
        V3 o{ro.x, ro.y, ro.z}, d{rd.x, rd.y, rd.z};
        const float4* tq = reinterpret_cast<const float4*>(sc.tris + tri);
        V3 ng = normalize(cross(load3(tq[1]), load3(tq[2])));
        in.p = o + d * hit.x;
        in.n = dot(d, ng) < 0.0f ? ng : -ng;
        in.wo = -d;
        in.rough = tri_type(sc.info[tri]) == kMatGGX ? sc.shade[tri].w : 0.0f;
      }
    }
    MlpActs act;
    mlp_inputs(m, in, lane, act.x0);
    f32x4_t o[3];
    mlp_forward(lw, act, o, lane);
    mlp_spill(hs[wave], o, lane);
    __syncthreads();
    MlpHead h = mlp_head(row, lane);
    if (valid) mlp_write_mix(h, reinterpret_cast<float*>(ps.mix + (size_t)i * kMlpMixF4), ps.mixc + i, lane);
    __syncthreads();
  }
}

// Training samples of the recorded paths (one thread per slot): the radiance that reached
// vertex d from its sampled direction is the sum of the contributions of the
// later hits divided by the throughput after the bounce (per channel, guarded
// as in the SD-tree record); the sample weight is avg_rgb(f |cos| / p_sample *
// Li).  Appended with one atomic per work-group and depth (the samples of a
// work-group's 16 waves combined in LDS: appends to one counter serialise);
// samples beyond the capacity are dropped.
constexpr uint32_t kEmitBlock = 1024;
__global__ void __launch_bounds__(kEmitBlock) k_mlp_emit(PathSoA ps, uint32_t max_depth, uint32_t max_vtx) {
  constexpr uint32_t kWaves = kEmitBlock / 64;
  __shared__ uint32_t s_n[kWaves + 1];
  __shared__ uint32_t s_max;
  uint32_t slot = blockIdx.x * kEmitBlock + threadIdx.x;
  uint32_t pid = slot * ps.mlp_stride + ps.mlp_phase;
  bool in = slot < ps.mlp_rec && pid < ps.n;
  uint32_t nh = in ? ps.nhit[pid] : 0u;
  uint32_t nv = nh > 0 ? nh - 1u : 0u;
  nv = nv < max_vtx ? nv : max_vtx;
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  // the deepest path of the work-group bounds its (uniform) depth loop
  uint32_t nh_max = nh;
  for (int off = 32; off >= 1; off >>= 1) nh_max = max(nh_max, (uint32_t)__shfl_xor((int)nh_max, off));
  if (threadIdx.x == 0) s_max = 0u;
  __syncthreads();
  if (lane == 0) atomicMax(&s_max, nh_max);
  __syncthreads();
  nh_max = s_max;
  const size_t st = (size_t)max_depth * ps.mlp_rec;
  float rx = 0.0f, ry = 0.0f, rz = 0.0f;
  for (int d = (int)nh_max - 2; d >= 0; --d) {
    if ((uint32_t)d + 1u < nh) {
      float4 b = ps.vB[(size_t)(d + 1) * ps.n + pid];
      rx += b.x;
      ry += b.y;
      rz += b.z;
    }
    bool ok = false;
    float w = 0.0f;
    size_t mi = (size_t)d * ps.mlp_rec + slot;
    float4 mc{0, 0, 0, 0};
    if ((uint32_t)d < nv && (uint32_t)d < ps.mlp_guide_depth) {
      mc = ps.mv[2 * st + mi];
      if (fbits(mc.w) & kVtxValid) {
        float4 va = ps.vA[(size_t)d * ps.n + pid];
        float4 bw = ps.mv[4 * st + mi];
        float pdf = va.w;
        bool fin = isfinite(rx) && rx >= 0.0f && isfinite(ry) && ry >= 0.0f && isfinite(rz) && rz >= 0.0f;
        if (fin && pdf > 0.0f) {
          float lx = va.x * pdf > 1e-4f ? rx / va.x : 0.0f;
          float ly = va.y * pdf > 1e-4f ? ry / va.y : 0.0f;
          float lz = va.z * pdf > 1e-4f ? rz / va.z : 0.0f;
          w = (bw.x * lx + bw.y * ly + bw.z * lz) / 3.0f;
          ok = isfinite(w) && w > 0.0f;
        }
      }
    }
    // work-group append: per-wave counts, one atomic
    uint64_t mask = __ballot(ok);
    if (lane == 0) s_n[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (uint32_t k = 0; k < kWaves; ++k) {
        uint32_t c = s_n[k];
        s_n[k] = t;
        t += c;
      }
      s_n[kWaves] = t ? atomicAdd(ps.mlp_count, t) : 0u;
    }
    __syncthreads();
    const uint32_t o_slot = s_n[kWaves] + s_n[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();  // s_n is rewritten by the next depth
    if (!ok || o_slot >= ps.mlp_cap) continue;
    float4 ma = ps.mv[mi], mb = ps.mv[st + mi], md = ps.mv[3 * st + mi];
    float4* o = ps.mlp_out + 4u * (size_t)o_slot;
    o[0] = ma;
    o[1] = mb;
    o[2] = make_float4(mc.x, mc.y, mc.z, w);
    o[3] = md;
  }
}

// Ray counters of a chunk: paths, radiance rays (live paths per depth) and
// shadow rays, summed over the queue regions and accumulated in 64 bits.
__global__ void __launch_bounds__(kQueueRegions) k_stats_add(const uint32_t* count, const uint32_t* scount,
                                                              uint32_t depth, uint32_t n,
                                                              unsigned long long* stats) {
  __shared__ unsigned long long part[2][kQueueRegions / 64];
  const uint32_t k = threadIdx.x;
  unsigned long long r = 0, sh = 0;
  for (uint32_t d = 0; d < depth; ++d) {
    r += count[d * kQueueRegions + k];
    sh += scount[d * kQueueRegions + k];
  }
  for (int off = 32; off >= 1; off >>= 1) {
    r += __shfl_xor(r, off);
    sh += __shfl_xor(sh, off);
  }
  if (lane_id() == 0) {
    part[0][k >> 6] = r;
    part[1][k >> 6] = sh;
  }
  __syncthreads();
  if (k == 0) {
    unsigned long long tr = 0, ts = 0;
    for (uint32_t w = 0; w < kQueueRegions / 64; ++w) {
      tr += part[0][w];
      ts += part[1][w];
    }
    stats[0] += n;
    stats[1] += tr;
    stats[2] += ts;
  }
}

// ------------------------------------------------------------------ host launchers
hipError_t launch_stats_add(const uint32_t* count, const uint32_t* scount, uint32_t depth, uint32_t n,
                            unsigned long long* stats, hipStream_t s) {
  hipLaunchKernelGGL(k_stats_add, dim3(1), dim3(kQueueRegions), 0, s, count, scount, depth, n, stats);
  return hipGetLastError();
}

static inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

hipError_t launch_trace_rays(const SceneView& sc, const float4* rays, uint32_t n, float4* hits, uint32_t* occl,
                             bool any, uint32_t* deep, uint32_t* err, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_trace_rays, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, sc, rays, n, hits, occl,
                     any ? 1u : 0u, deep, err);
  return hipGetLastError();
}

hipError_t launch_render_chunk(const RenderParams& rp, const Camera& cam, const SceneView& sc, PathSoA ps,
                               const SdTreeView& sd, const HashGridView& hg, const MortonView& mq, const MlpView& ml,
                               int guide, float* accum, hipStream_t s, hipEvent_t* ev) {
  // queue kernels cover the slots of every region in use
  const uint32_t nq = ps.nreg * ps.qreg;
  uint32_t nb = blocks_for(nq, kBlock);
  uint32_t ncnt = 2u * (rp.max_depth + 2u) * kQueueRegions;
  hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(ncnt, 64)), dim3(64), 0, s, ps.count, ncnt);
  hipLaunchKernelGGL(k_camera, dim3(blocks_for(ps.n, kBlock)), dim3(kBlock), 0, s, rp, cam, ps);
  // depth-0 counts of the regions: written by a kernel to stay stream ordered
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(kQueueRegions), 0, s, ps.count, ps.n, ps.qreg, ps.nreg);
  const bool record = guide != 0 && !(rp.flags & kFlagFinal);
  const bool count = ps.work != nullptr;
  // ev: per kernel class k in [0, kTimedKernels): pairs of events (begin, end)
  // per depth; class 0 extend, 1 shade, 2 shadow, 3 accumulate, 4 commit.
  auto mark = [&](int k, uint32_t d, int end) {
    if (ev) (void)hipEventRecord(ev[(k * kMaxDepthEv + d) * 2 + end], s);
  };
  for (uint32_t d = 0; d < rp.max_depth; ++d) {
    mark(0, d, 0);
    if (count) {
      hipLaunchKernelGGL(k_extend<true>, dim3(nb), dim3(kBlock), 0, s, sc, ps, d);
    } else {
      hipLaunchKernelGGL(k_extend<false>, dim3(nb), dim3(kBlock), 0, s, sc, ps, d);
    }
    mark(0, d, 1);
    mark(1, d, 0);
    const bool adjoint = ps.grad_alb != nullptr;
    const bool ris = (rp.flags & kFlagRis) != 0;
    const uint32_t nbs = blocks_for(nq, kShadeBlock);
    const bool tex = sc.tattr != nullptr || sc.vnorm != nullptr;
#define RTPG_SHADE_T(G, A, R)                                                                                       \
  if (tex) {                                                                                                        \
    hipLaunchKernelGGL((k_shade<G, A, R, true>), dim3(nbs), dim3(kShadeBlock), 0, s, rp, sc, ps, sd, hg, mq, d);  \
  } else {                                                                                                          \
    hipLaunchKernelGGL((k_shade<G, A, R, false>), dim3(nbs), dim3(kShadeBlock), 0, s, rp, sc, ps, sd, hg, mq, d); \
  }
#define RTPG_SHADE(G, A)    \
  if (ris) {                \
    RTPG_SHADE_T(G, A, true)  \
  } else {                  \
    RTPG_SHADE_T(G, A, false) \
  }
    if (adjoint) {
      // both passes of the adjoint run this one instantiation, so the replay
      // reproduces the primal paths operation for operation (RIS is refused)
      RTPG_SHADE_T(0, true, false)
    } else if (guide == 1) {
      RTPG_SHADE(1, false)
    } else if (guide == 2) {
      RTPG_SHADE(2, false)
    } else if (guide == 3) {
      // the learned guide's inference pass is timed with the shade class
      if ((rp.flags & kFlagGuideSample) && d < ps.mlp_guide_depth && !ps.mcache) {
        uint32_t ni = std::min(blocks_for(nq, 64u), kInferGrid);
        hipLaunchKernelGGL(k_mlp_infer, dim3(ni), dim3(kInferBlock), 0, s, sc, ps, ml, d);
      }
      RTPG_SHADE(3, false)
    } else {
      RTPG_SHADE(0, false)
    }
#undef RTPG_SHADE
#undef RTPG_SHADE_T
    mark(1, d, 1);
    mark(2, d, 0);
    if (adjoint) {
      hipLaunchKernelGGL((k_shadow<false, true>), dim3(nb), dim3(kBlock), 0, s, sc, ps, d, 0u);
    } else if (count) {
      hipLaunchKernelGGL(k_shadow<true>, dim3(nb), dim3(kBlock), 0, s, sc, ps, d, record ? 1u : 0u);
    } else {
      hipLaunchKernelGGL(k_shadow<false>, dim3(nb), dim3(kBlock), 0, s, sc, ps, d, record ? 1u : 0u);
    }
    mark(2, d, 1);
  }
  uint32_t npix = rp.width * rp.height;
  mark(3, 0, 0);
  hipLaunchKernelGGL(k_accum, dim3(blocks_for(npix, 256)), dim3(256), 0, s, ps, npix, rp.spp, accum);
  mark(3, 0, 1);
  mark(4, 0, 0);
  if (record && (guide == 1 || guide == 2 || (guide == 3 && sd.stree))) {
    if (guide != 2) {
      using Sh = CommitShape<1>;
      hipLaunchKernelGGL(k_commit<1>, dim3(blocks_for(ps.n, Sh::block * Sh::paths)), dim3(Sh::block), 0, s, ps,
                         rp.max_depth - 1u, sd, mq);
    } else {
      using Sh = CommitShape<2>;
      hipLaunchKernelGGL(k_commit<2>, dim3(blocks_for(ps.n, Sh::block * Sh::paths)), dim3(Sh::block), 0, s, ps,
                         rp.max_depth - 1u, sd, mq);
    }
  }
  if (record && guide == 3) {
    uint32_t ne = blocks_for(ps.mlp_rec, kEmitBlock);
    hipLaunchKernelGGL(k_mlp_emit, dim3(ne), dim3(kEmitBlock), 0, s, ps, rp.max_depth, rp.max_depth - 1u);
  }
  mark(4, 0, 1);
  return hipGetLastError();
}

}  // namespace rtpg
