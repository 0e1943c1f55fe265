/*
 * oracle.c — CPU oracle of the rtpg path (TEST INFRASTRUCTURE ONLY; see oracle.h).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so products and sums are
 * rounded one operation at a time, as the HIP kernels do where bit-exactness
 * is claimed).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define PI_F 3.14159265358979323846f
#define INV_PI_F 0.31830988618379067154f
#define TWO_PI_F 6.28318530717958647692f
#define INV_2PI_F 0.15915494309189533577f
#define LEAF_BIT 0x80000000u
#define INVALID 0xFFFFFFFFu

/* ======================================================================
 * random numbers
 * ==================================================================== */
typedef struct {
  uint64_t state, inc;
} pcg32;

static uint32_t pcg32_next(pcg32* r) {
  uint64_t old = r->state;
  r->state = old * 6364136223846793005ULL + r->inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
}

/* pcg32_srandom_r (O'Neill, pcg-c-basic) */
static pcg32 pcg32_seed(uint64_t initstate, uint64_t seq) {
  pcg32 r;
  r.state = 0u;
  r.inc = (seq << 1u) | 1u;
  pcg32_next(&r);
  r.state += initstate;
  pcg32_next(&r);
  return r;
}

static uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static pcg32 path_rng(uint64_t seed, uint32_t pixel, uint32_t sample) {
  return pcg32_seed(mix64(seed ^ mix64((uint64_t)sample)), (uint64_t)pixel);
}

/* to_uniform_float, Math/Math.h:682-695 */
float orc_to_uniform_float(uint32_t v) {
  uint32_t bits = (v >> 9) | 0x3F800000u;
  float f;
  memcpy(&f, &bits, 4);
  return f - 1.0f;
}

static float rnd(pcg32* r) { return orc_to_uniform_float(pcg32_next(r)); }

void orc_rng_stream(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t n, uint32_t* out) {
  pcg32 r = path_rng(seed, pixel, sample);
  for (uint32_t i = 0; i < n; ++i) out[i] = pcg32_next(&r);
}

void orc_pcg32_raw(uint64_t initstate, uint64_t seq, uint32_t n, uint32_t* out) {
  pcg32 r = pcg32_seed(initstate, seq);
  for (uint32_t i = 0; i < n; ++i) out[i] = pcg32_next(&r);
}

void orc_rng_floats(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t n, float* out) {
  pcg32 r = path_rng(seed, pixel, sample);
  for (uint32_t i = 0; i < n; ++i) out[i] = rnd(&r);
}

/* Xorshift32::next, Math/Random.h:44-50 */
uint32_t orc_xorshift32(uint32_t* state) {
  uint32_t y = *state;
  y = y ^ (y << 13);
  y = y ^ (y >> 17);
  return *state = y ^ (y << 5);
}

/* pcg1d, Math/Hash.h:13-18 */
uint32_t orc_pcg1d(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

/* xxhash, Math/Hash.h:50-66 */
uint32_t orc_xxhash32(uint32_t data, uint32_t seed) {
  const uint32_t P2 = 2246822519U, P3 = 3266489917U, P4 = 668265263U, P5 = 374761393U;
  uint32_t h = data * P3;
  h += seed + P5 + 4U;
  h = (h << 17) | (h >> 15);
  h *= P4;
  h ^= h >> 15;
  h *= P2;
  h ^= h >> 13;
  h *= P3;
  h ^= h >> 16;
  return h;
}

/* ======================================================================
 * Morton codes: bit-by-bit interleave (the product uses magic-number
 * spreads; this loop form is the definition).
 * ==================================================================== */
uint32_t orc_morton2(uint32_t x, uint32_t y) {
  uint32_t c = 0;
  for (int b = 0; b < 16; ++b) {
    c |= ((x >> b) & 1u) << (2 * b);
    c |= ((y >> b) & 1u) << (2 * b + 1);
  }
  return c;
}

uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t c = 0;
  for (int b = 0; b < 10; ++b) {
    c |= ((x >> b) & 1u) << (3 * b);
    c |= ((y >> b) & 1u) << (3 * b + 1);
    c |= ((z >> b) & 1u) << (3 * b + 2);
  }
  return c;
}

static void morton2_decode(uint32_t c, uint32_t* x, uint32_t* y) {
  uint32_t xx = 0, yy = 0;
  for (int b = 0; b < 16; ++b) {
    xx |= ((c >> (2 * b)) & 1u) << b;
    yy |= ((c >> (2 * b + 1)) & 1u) << b;
  }
  *x = xx;
  *y = yy;
}

/* ======================================================================
 * vectors
 * ==================================================================== */
typedef struct {
  float x, y, z;
} v3;
static v3 V(float x, float y, float z) {
  v3 r = {x, y, z};
  return r;
}
static v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3 scl(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static v3 divs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static float len(v3 a) { return sqrtf(dot(a, a)); }
static v3 nrm(v3 a) { return scl(a, 1.0f / sqrtf(dot(a, a))); }
static float avg3(v3 a) { return (a.x + a.y + a.z) / 3.0f; }
static float max3(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
static int finite3(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static v3 ld3(const float* p) { return V(p[0], p[1], p[2]); }

/* canonical maps, Math/VectorFunction.h:5698-5720 */
static void dir_to_canonical(v3 d, float* px, float* py) {
  if (!finite3(d)) {
    *px = 0.0f;
    *py = 0.0f;
    return;
  }
  float z = fminf(fmaxf(d.z, -1.0f), 1.0f);
  float phi = atan2f(d.y, d.x);
  while (phi < 0.0f) phi += TWO_PI_F;
  *px = (z + 1.0f) / 2.0f;
  *py = phi * INV_2PI_F;
}
static v3 canonical_to_dir(float px, float py) {
  float ct = 2.0f * px - 1.0f;
  float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
  float phi = TWO_PI_F * py;
  return V(st * cosf(phi), st * sinf(phi), ct);
}
void orc_dir_to_canonical(const float* d, float* uv) { dir_to_canonical(ld3(d), &uv[0], &uv[1]); }
void orc_canonical_to_dir(const float* uv, float* d) {
  v3 r = canonical_to_dir(uv[0], uv[1]);
  d[0] = r.x;
  d[1] = r.y;
  d[2] = r.z;
}

/* ONB, Math/VectorFunction.h:5666-5697 */
typedef struct {
  v3 t, b, n;
} onb;
static onb make_onb(v3 normal) {
  onb o;
  o.n = nrm(normal);
  if (fabsf(o.n.x) > fabsf(o.n.z)) {
    o.b = V(-o.n.y, o.n.x, 0.0f);
  } else {
    o.b = V(0.0f, -o.n.z, o.n.y);
  }
  o.b = nrm(o.b);
  o.t = nrm(cross(o.n, o.b));
  return o;
}
static v3 onb_local(const onb* o, v3 d) { return add(add(scl(o->t, d.x), scl(o->b, d.y)), scl(o->n, d.z)); }

/* ======================================================================
 * hash grid, Utils/OPX7UtilsGrid.h:157-353
 * ==================================================================== */
void orc_grid_cell_hash(const orc_grid* g, const float* pos, uint32_t* cell3, uint32_t* bucket, uint32_t* checksum) {
  uint32_t c[3];
  for (int k = 0; k < 3; ++k) {
    float f = (pos[k] - g->origin[k]) / g->size[k];
    float gk = (float)g->bounds[k] * f;
    /* CUDA's float->uint conversion saturates: negative -> 0 */
    uint32_t ci = gk > 0.0f ? (gk < 4294967040.0f ? (uint32_t)gk : 0xFFFFFFFFu) : 0u;
    c[k] = ci < g->bounds[k] - 1u ? ci : g->bounds[k] - 1u;
    cell3[k] = c[k];
  }
  uint32_t nb = g->slots / 32u;
  if (g->key_mode == 1u) {
    uint32_t key = orc_morton3(c[0], c[1], c[2]);
    *bucket = orc_pcg1d(key) % nb;
    *checksum = key + 1u;
  } else {
    *bucket = orc_pcg1d(g->slots + c[2] + orc_pcg1d(c[1] + orc_pcg1d(c[0]))) % nb;
    *checksum = 1u + g->bounds[0] * g->bounds[1] * c[2] + g->bounds[0] * c[1] + c[0];
  }
}

/* GetPrvCellIndex :298-319 (read-only) / GetCurCellIndex :266-297 (host branch) */
void orc_grid_find(const orc_grid* g, uint32_t* cur, const uint32_t* prv, const float* pos, uint32_t n, int insert,
                   uint32_t* out) {
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t c3[3], b, cs;
    orc_grid_cell_hash(g, pos + 3 * i, c3, &b, &cs);
    uint32_t res = INVALID;
    for (uint32_t k = 0; k < 32u; ++k) {
      uint32_t idx = 32u * b + k;
      if (insert) {
        uint32_t prev = cur[idx];
        if (prev == 0u) cur[idx] = cs;
        if (prev == 0u || prev == cs) {
          res = idx;
          break;
        }
      } else {
        uint32_t c = prv[idx];
        if (c == 0u) break;
        if (c == cs) {
          res = idx;
          break;
        }
      }
    }
    out[i] = res;
  }
}

/* One position's slot for the integrator: FindFromPrv (read-only probe of the
 * previous table, stops at an empty slot, GetPrvCellIndex :298-319) and
 * FindFromCur (insert with compare-and-swap 0 -> checksum, the device branch
 * of GetCurCellIndex :266-297; threads may insert concurrently). */
static uint32_t grid_find_prv1(const uint32_t* prv, uint32_t bucket, uint32_t cs) {
  for (uint32_t k = 0; k < 32u; ++k) {
    uint32_t c = prv[32u * bucket + k];
    if (c == 0u) return INVALID;
    if (c == cs) return 32u * bucket + k;
  }
  return INVALID;
}
static uint32_t grid_find_cur1(uint32_t* cur, uint32_t bucket, uint32_t cs) {
  for (uint32_t k = 0; k < 32u; ++k) {
    uint32_t expected = 0u;
    uint32_t idx = 32u * bucket + k;
    if (__atomic_compare_exchange_n(&cur[idx], &expected, cs, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return idx;
    if (expected == cs) return idx;
  }
  return INVALID;
}

/* ======================================================================
 * Morton quadtree, Utils/OPX7UtilsMorton.h:253-449, OPX7UtilsMortonBuild.cu
 * ==================================================================== */
static uint32_t level_offset(uint32_t level) { return ((1u << (2u * level)) - 1u) / 3u; }

static uint32_t canon_bin(float c, uint32_t nn) {
  uint32_t x = (uint32_t)(c * (float)nn);
  return x < nn ? x : nn - 1u;
}

/* MortonQuadTreeT::Record, leaf-only branch :262-285 */
void orc_mq_record(float* w, uint32_t per_cell, uint32_t level, const uint32_t* slot, const float* canon,
                   const float* val, uint32_t n) {
  uint32_t nn = 1u << level;
  for (uint32_t i = 0; i < n; ++i) {
    if (slot[i] == INVALID) continue;
    uint32_t code = orc_morton2(canon_bin(canon[2 * i], nn), canon_bin(canon[2 * i + 1], nn));
    w[(size_t)per_cell * slot[i] + level_offset(level) + code] += val[i];
  }
}

/* mortonBuildKernel :2-16, the serial loop with exact offsets (4^k-1)/3 */
void orc_mq_build(float* w, uint32_t level, uint32_t per_cell, uint32_t cells) {
  for (uint32_t c = 0; c < cells; ++c) {
    float* b = w + (size_t)per_cell * c;
    for (uint32_t i = 0; i < level; ++i) {
      uint32_t src = level_offset(level - i);
      uint32_t dst = level_offset(level - 1 - i);
      uint32_t ncode = 1u << (2u * (level - i));
      for (uint32_t code = 0; code < ncode; ++code) b[dst + (code >> 2)] += b[src + code];
    }
  }
}

/* MortonQuadTreeT::SampleAndPdf :286-344 on one pyramid b, drawing from r:
 * canonical (u, v) and the sphere density (pdf times 1/(4 pi), wrapper
 * :486-500).  An empty pyramid (root <= 0) samples the square uniformly. */
static float mq_sample1(const float* b, uint32_t level, pcg32* r, float* ou, float* ov) {
  float total = b[0];
  if (total <= 0.0f) {
    *ou = rnd(r);
    *ov = rnd(r);
    return 1.0f * 0.25f * INV_PI_F;
  }
  uint32_t size = 4, offset = 1, code = 0;
  for (uint32_t l = 0; l < level; ++l) {
    float wv[4];
    for (int k = 0; k < 4; ++k) wv[k] = b[offset + (code << 2) + (uint32_t)k];
    uint32_t local = 0;
    float partial = wv[0] + wv[2];
    float boundary = partial / total;
    float sample = rnd(r);
    if (sample < boundary) {
      sample /= boundary;
      boundary = wv[0] / partial;
    } else {
      partial = wv[1] + wv[3];
      sample = (sample - boundary) / (1.0f - boundary);
      boundary = wv[1] / partial;
      local |= 1u;
    }
    if (sample >= boundary) local |= 2u;
    total = wv[local];
    code = (code << 2) + local;
    offset += size;
    size *= 4u;
  }
  uint32_t xi, yi;
  morton2_decode(code, &xi, &yi);
  float nn = (float)(1u << level);
  float a = rnd(r);
  float c = rnd(r);
  *ou = ((float)xi + a) / nn;
  *ov = ((float)yi + c) / nn;
  return total / b[0] * (nn * nn) * 0.25f * INV_PI_F;
}

void orc_mq_sample(const float* w, uint32_t per_cell, uint32_t level, const uint32_t* slot, uint32_t n,
                   uint64_t seed, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    pcg32 r = pcg32_seed(seed, i);
    out[3 * i + 2] = mq_sample1(w + (size_t)per_cell * slot[i], level, &r, &out[3 * i], &out[3 * i + 1]);
  }
}

static float mq_pdf1(const float* b, uint32_t level, float cu, float cv) {
  if (b[0] <= 0.0f) return 1.0f * 0.25f * INV_PI_F;
  uint32_t nn = 1u << level;
  uint32_t code = orc_morton2(canon_bin(cu, nn), canon_bin(cv, nn));
  return b[level_offset(level) + code] / b[0] * (float)(nn * nn) * 0.25f * INV_PI_F;
}

/* MortonQuadTreeT::Pdf :401-414 */
void orc_mq_pdf(const float* w, uint32_t per_cell, uint32_t level, const uint32_t* slot, const float* canon,
                uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i)
    out[i] = mq_pdf1(w + (size_t)per_cell * slot[i], level, canon[2 * i], canon[2 * i + 1]);
}

/* ======================================================================
 * SD-tree, Utils/OPX7UtilsPathGuiding.h
 * ==================================================================== */
static int child_idx(float* px, float* py) { /* DTreeNodeT::GetChildIdx :69-82 */
  int r = 0;
  if (*px < 0.5f) {
    *px *= 2.0f;
  } else {
    *px = 2.0f * *px - 1.0f;
    r |= 1;
  }
  if (*py < 0.5f) {
    *py *= 2.0f;
  } else {
    *py = 2.0f * *py - 1.0f;
    r |= 2;
  }
  return r;
}

/* DTreeNodeT::GetArea :342-355 (recursive) */
static float dnode_area(const float* sums, const uint16_t* child, uint32_t k) {
  float result = 0.0f;
  for (int i = 0; i < 4; ++i) {
    if (sums[4 * k + i] > 0.0f) {
      uint16_t c = child[4 * k + i];
      if (c == 0) {
        result += 1.0f / 4.0f;
      } else {
        result += dnode_area(sums, child, c) / 4.0f;
      }
    }
  }
  return result;
}

/* DTreeNodeT::Build :356-369 (recursive post-order) */
static void dnode_build(float* sums, const uint16_t* child, uint32_t k) {
  for (int i = 0; i < 4; ++i) {
    uint16_t c = child[4 * k + i];
    if (c == 0) continue;
    dnode_build(sums, child, c);
    float sum = 0.0f;
    for (int j = 0; j < 4; ++j) sum += sums[4 * c + j];
    sums[4 * k + i] = sum;
  }
}

/* RTDTreeT::Build :948-957, per DTree of the concatenated pool */
void orc_dtree_build(uint32_t n_dtrees, const uint32_t* cnt, float* sums, const uint16_t* child, float* out_sum,
                     float* out_area) {
  size_t base = 0;
  for (uint32_t j = 0; j < n_dtrees; ++j) {
    float* s = sums + 4 * base;
    const uint16_t* c = child + 4 * base;
    dnode_build(s, c, 0);
    out_area[j] = dnode_area(s, c, 0);
    float sum = 0.0f;
    for (int i = 0; i < 4; ++i) sum += s[i];
    out_sum[j] = sum;
    base += cnt[j];
  }
}

/* RTDTreeT::Reset :888-947 — DFS with an explicit LIFO stack; a source leaf
 * continues as a node of the new tree whose four sums are a quarter of the slot. */
uint32_t orc_dtree_reset(const float* src_sums, const uint16_t* src_child, float total, uint32_t max_depth,
                         float th, uint16_t* out_child, uint32_t* depth_out) {
  enum { CAP = 65536 };
  typedef struct {
    uint32_t dst, src, depth, self;
    float val;
  } ent;
  ent* st = (ent*)malloc(sizeof(ent) * 4 * CAP);
  size_t sp = 0;
  uint32_t count = 1, maxd = 0;
  memset(out_child, 0, 4 * sizeof(uint16_t));
  st[sp++] = (ent){0, 0, 1, 0, 0.0f};
  int stop = 0;
  while (sp > 0 && !stop) {
    ent e = st[--sp];
    if (e.depth > maxd) maxd = e.depth;
    for (int i = 0; i < 4; ++i) {
      float si = e.self ? e.val : src_sums[4 * e.src + i];
      double fraction = total > 0.0f ? (double)(si / total) : pow(0.25, (double)e.depth);
      if (e.depth < max_depth && fraction > (double)th) {
        uint16_t sch = e.self ? 0 : src_child[4 * e.src + i];
        ent ne;
        ne.dst = count;
        ne.depth = e.depth + 1;
        if (!e.self && sch != 0) {
          ne.src = sch;
          ne.self = 0;
          ne.val = 0.0f;
        } else {
          ne.src = 0;
          ne.self = 1;
          ne.val = si / 4.0f;
        }
        st[sp++] = ne;
        out_child[4 * e.dst + i] = (uint16_t)count;
        memset(out_child + 4 * (size_t)count, 0, 4 * sizeof(uint16_t));
        count++;
        if (count > 65535u) {
          stop = 1;
          break;
        }
      }
    }
  }
  free(st);
  *depth_out = maxd;
  return count;
}

uint64_t orc_split_threshold(int mode, float c, int iter, int spp) {
  if (mode == 0) return (uint64_t)(sqrt(pow(2.0, iter) * spp / 4.0f) * (double)c);
  return (uint64_t)((double)c * sqrt(pow(2.0, iter)));
}

/* RTSTreeT::Refine :1275-1311 with ShallSplit :1270-1274 and SubDivide
 * :1231-1247; octree != 0: RTSTree2T::Refine :2108-2142 with SubDivide
 * :2062-2077 (8 children, weight / 8), the octree node stored as three binary
 * levels (pairs at n, n+2..n+5, leaves n+6..n+13). */
uint32_t orc_stree_refine(const uint32_t* stree, uint32_t n_nodes, const float* b_sw, const uint32_t* b_cnt,
                          const uint32_t* s_cnt, uint64_t th, int max_mb, uint32_t* out, uint32_t cap, uint32_t* map,
                          uint32_t* n_dtrees_out, int octree) {
  typedef struct {
    int leaf;
    uint32_t child;
    float sw;
    int64_t src;
  } hn;
  if (cap < n_nodes) return 0;
  hn* h = (hn*)malloc(sizeof(hn) * cap);
  uint32_t n = n_nodes;
  for (uint32_t i = 0; i < n; ++i) {
    h[i].leaf = (stree[i] & LEAF_BIT) != 0;
    h[i].child = h[i].leaf ? 0 : stree[i];
    h[i].src = h[i].leaf ? (int64_t)(stree[i] & ~LEAF_BIT) : -1;
    h[i].sw = h[i].leaf ? b_sw[h[i].src] : 0.0f;
  }
  int refine = 1;
  if (max_mb >= 0) {
    size_t fp = 0;
    for (uint32_t i = 0; i < n; ++i)
      fp += h[i].leaf ? 24u * ((size_t)b_cnt[h[i].src] + s_cnt[h[i].src]) + 48u : 96u;
    if (fp / 1000000 >= (size_t)max_mb) refine = 0;
  }
  if (refine) {
    size_t scap = 2 * (size_t)cap + 2, sp = 0;
    uint32_t* st = (uint32_t*)malloc(sizeof(uint32_t) * scap);
    st[sp++] = 0;
    while (sp > 0) {
      uint32_t idx = st[--sp];
      if (octree && h[idx].leaf && (uint64_t)n < 0xFFFFFFFFull - 1 && h[idx].sw > (float)th) {
        if (n + 14 > cap) {
          free(st);
          free(h);
          return 0;
        }
        float sw8 = h[idx].sw / 8.0f;
        int64_t src = h[idx].src;
        h[idx].leaf = 0;
        h[idx].child = n;
        h[idx].src = -1;
        for (uint32_t k = 0; k < 6; ++k) { /* the x pair's y pairs, then their z pairs */
          h[n + k].leaf = 0;
          h[n + k].child = n + 2 + 2 * k;
          h[n + k].src = -1;
          h[n + k].sw = 0.0f;
        }
        for (uint32_t k = 6; k < 14; ++k) {
          h[n + k].leaf = 1;
          h[n + k].child = 0;
          h[n + k].sw = sw8;
          h[n + k].src = src;
        }
        n += 14;
      } else if (h[idx].leaf && (uint64_t)n < 0xFFFFFFFFull - 1 && h[idx].sw > (float)th) {
        if (n + 2 > cap) {
          free(st);
          free(h);
          return 0;
        }
        for (int i = 0; i < 2; ++i) {
          h[n + i].leaf = 1;
          h[n + i].child = 0;
          h[n + i].sw = h[idx].sw / 2.0f;
          h[n + i].src = h[idx].src;
        }
        h[idx].leaf = 0;
        h[idx].child = n;
        h[idx].src = -1;
        n += 2;
      }
      if (!h[idx].leaf) {
        st[sp++] = h[idx].child;
        st[sp++] = h[idx].child + 1;
      }
    }
    free(st);
  }
  uint32_t nd = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (h[i].leaf) {
      out[i] = LEAF_BIT | nd;
      map[nd++] = (uint32_t)h[i].src;
    } else {
      out[i] = h[i].child;
    }
  }
  *n_dtrees_out = nd;
  free(h);
  return n;
}

/* STreeT::GetDTreeWrapper :639-648, descent :561-567 (child index clamped) */
static uint32_t stree_lookup1(const uint32_t* stree, const float* lo, float size, v3 p) {
  float q[3] = {(p.x - lo[0]) / size, (p.y - lo[1]) / size, (p.z - lo[2]) / size};
  uint32_t w = stree[0];
  int axis = 0;
  while (!(w & LEAF_BIT)) {
    float pa = q[axis] * 2.0f;
    int idx = (int)pa;
    idx = idx < 0 ? 0 : (idx > 1 ? 1 : idx);
    q[axis] = pa - (float)idx;
    w = stree[w + (uint32_t)idx];
    axis = axis == 2 ? 0 : axis + 1;
  }
  return w & ~LEAF_BIT;
}

void orc_stree_lookup(const uint32_t* stree, const float* aabb_min, float size, const float* pos, uint32_t n,
                      uint32_t* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = stree_lookup1(stree, aabb_min, size, ld3(pos + 3 * i));
}

static void atomic_addf(float* p, float v) {
  uint32_t* u = (uint32_t*)p;
  uint32_t old = __atomic_load_n(u, __ATOMIC_RELAXED), nw;
  for (;;) {
    float f;
    memcpy(&f, &old, 4);
    f += v;
    memcpy(&nw, &f, 4);
    if (__atomic_compare_exchange_n(u, &old, nw, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
  }
}

/* DTreeT::Impl<Nearest>::RecordIrradiance :419-427 + DTreeNodeT::Record :113-125 */
static void dtree_record1(float* b_sw, float* b_sums, const uint16_t* b_child, uint32_t base, uint32_t cell,
                          float cu, float cv, float irr, float sw) {
  if (!(isfinite(sw) && sw > 0.0f)) return;
  atomic_addf(&b_sw[cell], sw);
  if (!(isfinite(irr) && irr > 0.0f)) return;
  uint32_t node = 0;
  int idx = child_idx(&cu, &cv);
  for (;;) {
    uint16_t c = b_child[4 * ((size_t)base + node) + idx];
    if (c == 0) break;
    node = c;
    idx = child_idx(&cu, &cv);
  }
  atomic_addf(&b_sums[4 * ((size_t)base + node) + idx], irr * sw);
}

/* ---- recording filters ------------------------------------------------
 * Stochastic form (what the GPU library runs): the record moves to a uniform
 * point of the reference's box and is recorded there, so each node's expected
 * weight and flux equal the reference's box-filtered values.  Its own stream
 * per (seed, pixel, sample, vertex depth), five draws: three for the spatial
 * point, two for the directional one. */
enum { FILTER_SPATIAL = 1, FILTER_DIRECTIONAL = 2 };

static pcg32 filter_rng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t depth) {
  return pcg32_seed(mix64(seed ^ 0xD6E8FEB86659FD93ULL ^ mix64(((uint64_t)sample << 8) | depth)), (uint64_t)pixel);
}

/* STreeT::GetDTreeWrapper(p, size) :639-648 with the descent of :556-567,
 * also returning the depth of the leaf (levels below the root). */
static uint32_t stree_lookup_depth(const uint32_t* stree, const float* lo, float size, v3 p, uint32_t* depth) {
  float q[3] = {(p.x - lo[0]) / size, (p.y - lo[1]) / size, (p.z - lo[2]) / size};
  uint32_t w = stree[0], d = 0;
  int axis = 0;
  while (!(w & LEAF_BIT)) {
    float pa = q[axis] * 2.0f;
    int idx = (int)pa;
    idx = idx < 0 ? 0 : (idx > 1 ? 1 : idx);
    q[axis] = pa - (float)idx;
    w = stree[w + (uint32_t)idx];
    axis = axis == 2 ? 0 : axis + 1;
    ++d;
  }
  *depth = d;
  return w & ~LEAF_BIT;
}

/* Spatial box (TraceVertexT::Impl<SpatialFilterBox> :806-813, STreeT::Record
 * :649-655): the box is the voxel of the leaf holding p (GetDTreeWrapper's
 * size, halved once per level along the level's axis) centred on p; returns
 * the leaf holding p + (u - 1/2) * voxel, INVALID outside the tree's cube. */
static uint32_t stree_box_jitter1(const uint32_t* stree, const float* lo, float size, v3 p, float u0, float u1,
                                  float u2) {
  uint32_t d;
  (void)stree_lookup_depth(stree, lo, size, p, &d);
  float e[3] = {size, size, size};
  for (uint32_t k = 0; k < d; ++k) e[k % 3] *= 0.5f;
  v3 q = V(p.x + (u0 - 0.5f) * e[0], p.y + (u1 - 0.5f) * e[1], p.z + (u2 - 0.5f) * e[2]);
  float qx = (q.x - lo[0]) / size, qy = (q.y - lo[1]) / size, qz = (q.z - lo[2]) / size;
  if (!(qx >= 0.0f && qx < 1.0f && qy >= 0.0f && qy < 1.0f && qz >= 0.0f && qz < 1.0f)) return INVALID;
  return stree_lookup1(stree, lo, size, q);
}

/* DTreeNodeT::GetDepth :322-333 on the building tree rooted at `base`. */
static int dtree_depth1(const uint16_t* b_child, uint32_t base, float cu, float cv) {
  uint32_t node = 0;
  int depth = 1;
  int idx = child_idx(&cu, &cv);
  for (;;) {
    uint16_t c = b_child[4 * ((size_t)base + node) + idx];
    if (c == 0) break;
    node = c;
    idx = child_idx(&cu, &cv);
    ++depth;
  }
  return depth;
}

/* Directional box (DTreeT::Impl<DirectionalFilterBox> :429-443): (u, v)
 * moves by (r - 1/2) 0.5^depth; 0 when it leaves [0,1)^2 (flux dropped). */
static int dtree_box_jitter1(const uint16_t* b_child, uint32_t base, float* cu, float* cv, float r0, float r1) {
  int depth = dtree_depth1(b_child, base, *cu, *cv);
  float e = 1.0f;
  for (int k = 0; k < depth; ++k) e *= 0.5f;
  *cu = *cu + (r0 - 0.5f) * e;
  *cv = *cv + (r1 - 0.5f) * e;
  return *cu >= 0.0f && *cu < 1.0f && *cv >= 0.0f && *cv < 1.0f;
}

/* filter: FILTER_* bits; the draws of record i come from
 * filter_rng(seed, i, batch, 0). */
void orc_sdtree_record(const uint32_t* stree, const float* aabb_min, float size, const uint32_t* b_base,
                       float* b_sw, float* b_sums, const uint16_t* b_child, const float* pos, const float* canon,
                       const float* irr, const float* sw, uint32_t n, uint32_t filter, uint64_t seed,
                       uint32_t batch) {
  for (uint32_t i = 0; i < n; ++i) {
    v3 p = ld3(pos + 3 * i);
    uint32_t cell = stree_lookup1(stree, aabb_min, size, p);
    float cu = canon[2 * i], cv = canon[2 * i + 1], val = irr[i];
    if (filter) {
      pcg32 fr = filter_rng(seed, i, batch, 0);
      float f[5];
      for (int k = 0; k < 5; ++k) f[k] = rnd(&fr);
      if (filter & FILTER_SPATIAL) cell = stree_box_jitter1(stree, aabb_min, size, p, f[0], f[1], f[2]);
      if (cell == INVALID) continue;
      if ((filter & FILTER_DIRECTIONAL) && !dtree_box_jitter1(b_child, b_base[cell], &cu, &cv, f[3], f[4])) val = 0.0f;
    }
    dtree_record1(b_sw, b_sums, b_child, b_base[cell], cell, cu, cv, val, sw[i]);
  }
}

/* Deterministic box filters exactly as the reference states them (not run on
 * the GPU; the reference to which the stochastic form is checked in
 * expectation).  DTreeNodeT::Record(origin, size, nodeOrigin, nodeSize,
 * value) :126-160: every leaf slot overlapped by the square [origin,
 * origin + size]^2 receives value times the overlap area. */
static void dtree_record_square(float* b_sums, const uint16_t* b_child, uint32_t base, uint32_t node, float ox,
                                float oy, float size, float nox, float noy, float nsize, float value) {
  float cs = nsize / 2.0f;
  for (int i = 0; i < 4; ++i) {
    float cx = nox + ((i & 1) ? cs : 0.0f), cy = noy + ((i & 2) ? cs : 0.0f);
    float lx = fmaxf(fminf(ox + size, cx + cs) - fmaxf(ox, cx), 0.0f);
    float ly = fmaxf(fminf(oy + size, cy + cs) - fmaxf(oy, cy), 0.0f);
    float w = lx * ly;
    if (w > 0.0f) {
      uint16_t c = b_child[4 * ((size_t)base + node) + i];
      if (c == 0) {
        b_sums[4 * ((size_t)base + node) + i] += value * w;
      } else {
        dtree_record_square(b_sums, b_child, base, c, ox, oy, size, cx, cy, cs, value);
      }
    }
  }
}

/* DTreeT::Impl<dFilter>::RecordIrradiance :416-443 into DTree `cell`. */
static void dtree_record_filtered(float* b_sw, float* b_sums, const uint16_t* b_child, uint32_t base,
                                  uint32_t cell, float cu, float cv, float irr, float sw, int box) {
  if (!box) {
    dtree_record1(b_sw, b_sums, b_child, base, cell, cu, cv, irr, sw);
    return;
  }
  if (!(isfinite(sw) && sw > 0.0f)) return;
  b_sw[cell] += sw;
  if (!(isfinite(irr) && irr > 0.0f)) return;
  int depth = dtree_depth1(b_child, base, cu, cv);
  float e = powf(0.5f, (float)depth);
  dtree_record_square(b_sums, b_child, base, 0, cu - e / 2.0f, cv - e / 2.0f, e, 0.0f, 0.0f, 1.0f,
                      irr * sw / (e * e));
}

/* STreeNodeT::Record :573-611: every leaf whose voxel [mn, mn + sz] overlaps
 * the box [b0, b1] records with statistical weight sw times the overlap
 * volume (sw already divided by the box volume, STreeT::Record :649-655). */
static void stree_record_box(const uint32_t* stree, uint32_t word, int axis, const float* mn, const float* sz,
                             const float* b0, const float* b1, const uint32_t* b_base, float* b_sw, float* b_sums,
                             const uint16_t* b_child, float cu, float cv, float irr, float sw, int dbox) {
  float w = 1.0f;
  for (int k = 0; k < 3; ++k) w *= fmaxf(fminf(b1[k], mn[k] + sz[k]) - fmaxf(b0[k], mn[k]), 0.0f);
  if (!(w > 0.0f)) return;
  if (word & LEAF_BIT) {
    uint32_t cell = word & ~LEAF_BIT;
    dtree_record_filtered(b_sw, b_sums, b_child, b_base[cell], cell, cu, cv, irr, sw * w, dbox);
    return;
  }
  float csz[3] = {sz[0], sz[1], sz[2]};
  csz[axis] /= 2.0f;
  for (int i = 0; i < 2; ++i) {
    float cmn[3] = {mn[0], mn[1], mn[2]};
    if (i) cmn[axis] += csz[axis];
    stree_record_box(stree, stree[word + (uint32_t)i], axis == 2 ? 0 : axis + 1, cmn, csz, b0, b1, b_base, b_sw,
                     b_sums, b_child, cu, cv, irr, sw, dbox);
  }
}

/* The reference's filtered record (TraceVertexT::Impl<sFilter>::Record
 * :796-813 -> DTreeWrapperT::Record -> DTreeT::Impl<dFilter>) of n records;
 * spatial, directional: 0 Nearest, 1 Box. */
void orc_sdtree_record_box(const uint32_t* stree, const float* aabb_min, float size, const uint32_t* b_base,
                           float* b_sw, float* b_sums, const uint16_t* b_child, const float* pos,
                           const float* canon, const float* irr, const float* sw, uint32_t n, int spatial,
                           int directional) {
  for (uint32_t i = 0; i < n; ++i) {
    v3 p = ld3(pos + 3 * i);
    uint32_t d;
    uint32_t cell = stree_lookup_depth(stree, aabb_min, size, p, &d);
    if (!spatial) {
      dtree_record_filtered(b_sw, b_sums, b_child, b_base[cell], cell, canon[2 * i], canon[2 * i + 1], irr[i],
                            sw[i], directional);
      continue;
    }
    float e[3] = {size, size, size};
    for (uint32_t k = 0; k < d; ++k) e[k % 3] /= 2.0f;
    float pc[3] = {p.x, p.y, p.z};
    float b0[3], b1[3], mn[3] = {aabb_min[0], aabb_min[1], aabb_min[2]}, sz[3] = {size, size, size};
    for (int k = 0; k < 3; ++k) {
      b0[k] = pc[k] - e[k] * 0.5f;
      b1[k] = pc[k] + e[k] * 0.5f;
    }
    float volume = e[0] * e[1] * e[2];
    stree_record_box(stree, stree[0], 0, mn, sz, b0, b1, b_base, b_sw, b_sums, b_child, canon[2 * i],
                     canon[2 * i + 1], irr[i], sw[i] / volume, directional);
  }
}

static float dtree_mean(float sum, float sw) { /* DTreeT::GetMean :444-448 */
  if (sw <= 0.0f) return 0.0f;
  float factor = INV_2PI_F / (2.0f * sw);
  return factor * sum;
}

/* DTreeNodeT::Sample :228-277 + clamp of DTreeT::Sample :477-482; *pdf: the
 * sphere density of the drawn point accumulated along the same descent, as
 * DTreeNodeT::SampleAndPdf :178-227 does (4 s_i / sum per level, 1 / (4 pi)). */
static void dtree_sample1(const float* sums, const uint16_t* child, float sum, float sw, pcg32* r, float* ox,
                          float* oy, float* pdf) {
  if (dtree_mean(sum, sw) <= 0.0f) {
    *ox = rnd(r);
    *oy = rnd(r);
    *pdf = 1.0f * INV_2PI_F / 2.0f;
    return;
  }
  float result = 1.0f;
  uint32_t node = 0;
  float rx = 0.0f, ry = 0.0f, size = 1.0f;
  for (;;) {
    const float* s = sums + 4 * node;
    int idx = 0;
    float partial = s[0] + s[2];
    float total = s[0] + s[1] + s[2] + s[3];
    float boundary = partial / total;
    float ox0 = 0.0f, oy0 = 0.0f;
    float sample = rnd(r);
    if (sample < boundary) {
      sample = sample / boundary;
      boundary = s[0] / partial;
    } else {
      partial = total - partial;
      ox0 = 0.5f;
      sample = (sample - boundary) / (1.0f - boundary);
      boundary = s[1] / partial;
      idx |= 1;
    }
    if (sample >= boundary) {
      oy0 = 0.5f;
      idx |= 2;
    }
    rx = rx + size * ox0;
    ry = ry + size * oy0;
    size *= 0.5f;
    result = s[idx] > 0.0f ? result * (4.0f * s[idx] / total) : 0.0f;
    uint16_t c = child[4 * node + idx];
    if (c == 0 || s[idx] <= 0.0f) {
      float a = rnd(r);
      float b = rnd(r);
      rx = rx + size * a;
      ry = ry + size * b;
      break;
    }
    node = c;
  }
  *ox = fminf(fmaxf(rx, 0.0f), 1.0f);
  *oy = fminf(fmaxf(ry, 0.0f), 1.0f);
  *pdf = result * INV_2PI_F / 2.0f;
}

/* DTreeNodeT::Pdf :278-316 times 1/(4 pi) (DTreeT::Pdf :484-489) */
static float dtree_pdf1(const float* sums, const uint16_t* child, float sum, float sw, float px, float py) {
  if (dtree_mean(sum, sw) <= 0.0f) return 1.0f * INV_2PI_F / 2.0f;
  float result = 1.0f;
  uint32_t node = 0;
  int idx = child_idx(&px, &py);
  for (;;) {
    const float* s = sums + 4 * node;
    float total = s[0] + s[1] + s[2] + s[3];
    if (s[idx] <= 0.0f) {
      result = 0.0f;
      break;
    }
    result *= 4.0f * s[idx] / total;
    uint16_t c = child[4 * node + idx];
    if (c == 0) break;
    node = c;
    idx = child_idx(&px, &py);
  }
  return result * INV_2PI_F / 2.0f;
}

void orc_sdtree_sample(const uint32_t* s_base, const float* s_sum, const float* s_sw, const float* s_sums,
                       const uint16_t* s_child, const uint32_t* cells, uint32_t n, uint64_t seed, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    pcg32 r = pcg32_seed(seed, i);
    uint32_t c = cells[i];
    float pdf;
    dtree_sample1(s_sums + 4 * (size_t)s_base[c], s_child + 4 * (size_t)s_base[c], s_sum[c], s_sw[c], &r,
                  &out[2 * i], &out[2 * i + 1], &pdf);
  }
}

void orc_sdtree_pdf(const uint32_t* s_base, const float* s_sum, const float* s_sw, const float* s_sums,
                    const uint16_t* s_child, const uint32_t* cells, const float* canon, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t c = cells[i];
    out[i] = dtree_pdf1(s_sums + 4 * (size_t)s_base[c], s_child + 4 * (size_t)s_base[c], s_sum[c], s_sw[c],
                        canon[2 * i], canon[2 * i + 1]);
  }
}

/* ======================================================================
 * geometry: Moller-Trumbore, two-sided, operation for operation as the HIP
 * kernel (render.hip isect_tri): the last product of every cross and dot
 * product fused (fmaf, correctly rounded here as on the GPU), barycentric
 * tests on u det and v det against det, one division for a passing triangle.
 * The reference's triangle test is OptiX's (closed): parity with it is
 * unpinned; brute force over this test pins the traversal.
 * ==================================================================== */
int orc_isect_tri(const float* o, const float* d, const float* a, const float* e1, const float* e2, float tmin,
                  float* tbest, float* ou, float* ov) {
  float px = fmaf(d[1], e2[2], -(d[2] * e2[1]));
  float py = fmaf(d[2], e2[0], -(d[0] * e2[2]));
  float pz = fmaf(d[0], e2[1], -(d[1] * e2[0]));
  float det = fmaf(e1[0], px, fmaf(e1[1], py, e1[2] * pz));
  if (det == 0.0f) return 0;
  float sx = o[0] - a[0], sy = o[1] - a[1], sz = o[2] - a[2];
  float ud = fmaf(sx, px, fmaf(sy, py, sz * pz));
  if (det > 0.0f ? !(ud >= 0.0f && ud <= det) : !(ud <= 0.0f && ud >= det)) return 0;
  float qx = fmaf(sy, e1[2], -(sz * e1[1]));
  float qy = fmaf(sz, e1[0], -(sx * e1[2]));
  float qz = fmaf(sx, e1[1], -(sy * e1[0]));
  float vd = fmaf(d[0], qx, fmaf(d[1], qy, d[2] * qz));
  float uvd = ud + vd;
  if (det > 0.0f ? !(vd >= 0.0f && uvd <= det) : !(vd <= 0.0f && uvd >= det)) return 0;
  float inv = 1.0f / det;
  float t = fmaf(e2[0], qx, fmaf(e2[1], qy, e2[2] * qz)) * inv;
  if (!(t > tmin && t < *tbest)) return 0;
  *tbest = t;
  *ou = ud * inv;
  *ov = vd * inv;
  return 1;
}

void orc_brute_closest(const float* verts, const uint32_t* idx, uint32_t nt, const float* rays, uint32_t n,
                       float* hits) {
  for (uint32_t r = 0; r < n; ++r) {
    const float* ray = rays + 8 * r;
    float tb = ray[7], u = 0, v = 0;
    uint32_t best = INVALID;
    for (uint32_t t = 0; t < nt; ++t) {
      const float* a = verts + 3 * idx[3 * t];
      const float* b = verts + 3 * idx[3 * t + 1];
      const float* c = verts + 3 * idx[3 * t + 2];
      float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      if (orc_isect_tri(ray, ray + 4, a, e1, e2, ray[3], &tb, &u, &v)) best = t;
    }
    hits[4 * r] = best == INVALID ? INFINITY : tb;
    hits[4 * r + 1] = u;
    hits[4 * r + 2] = v;
    memcpy(&hits[4 * r + 3], &best, 4);
  }
}

/* ======================================================================
 * scene + binary BVH (oracle-owned acceleration, binned SAH)
 * ==================================================================== */
typedef struct {
  float lo[3], hi[3];
  int32_t left;   /* >= 0 inner (right = left + 1); < 0 leaf */
  uint32_t first, count;
} bnode;

struct orc_scene {
  uint32_t nt;
  float* tri;      /* 9 floats per triangle: v0, e1, e2 (original order) */
  float* albedo;   /* 4 per tri */
  float* spec;     /* 4 per tri or NULL */
  uint32_t* mat;
  int32_t* light_of;
  uint32_t nl;
  float* light;    /* per light: v0(3) e1(3) e2(3) area, pdf, Le(3) = 15 floats */
  float* cdf;
  bnode* nodes;
  uint32_t nn;
  uint32_t* prim;  /* leaf order -> triangle */
  uint32_t* light_tri; /* light -> triangle */
  /* optional attributes (original triangle order), NULL when absent */
  float* tc;          /* 6 per triangle: u, v at vertices 0, 1, 2 */
  uint32_t* tid;      /* 4 per triangle: diffuse, specular, emission on hit, emission light-sampled */
  float* vn;          /* 9 per triangle: vertex normals */
  uint32_t* texels;   /* RGBA8 */
  uint32_t* tex;      /* 4 per texture: first texel, width, height, wrap */
};

typedef struct {
  float lo[3], hi[3], c[3];
} pref;

static float box_area(const float* lo, const float* hi) {
  float d0 = fmaxf(hi[0] - lo[0], 0.0f), d1 = fmaxf(hi[1] - lo[1], 0.0f), d2 = fmaxf(hi[2] - lo[2], 0.0f);
  return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
}

static void bvh_build(orc_scene* s) {
  uint32_t nt = s->nt;
  pref* r = (pref*)malloc(sizeof(pref) * nt);
  s->prim = (uint32_t*)malloc(sizeof(uint32_t) * nt);
  for (uint32_t t = 0; t < nt; ++t) {
    const float* q = s->tri + 9 * t;
    for (int k = 0; k < 3; ++k) {
      float a = q[k], b = q[k] + q[3 + k], c = q[k] + q[6 + k];
      r[t].lo[k] = fminf(a, fminf(b, c));
      r[t].hi[k] = fmaxf(a, fmaxf(b, c));
      r[t].c[k] = 0.5f * (r[t].lo[k] + r[t].hi[k]);
    }
    s->prim[t] = t;
  }
  size_t cap = 2 * (size_t)nt + 1;
  s->nodes = (bnode*)malloc(sizeof(bnode) * cap);
  s->nn = 1;
  s->nodes[0].first = 0;
  s->nodes[0].count = nt;
  uint32_t* st = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  size_t sp = 0;
  st[sp++] = 0;
  enum { NB = 16 };
  while (sp) {
    uint32_t ni = st[--sp];
    bnode* n = &s->nodes[ni];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = n->first; i < n->first + n->count; ++i) {
      pref* p = &r[s->prim[i]];
      for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(lo[k], p->lo[k]);
        hi[k] = fmaxf(hi[k], p->hi[k]);
        clo[k] = fminf(clo[k], p->c[k]);
        chi[k] = fmaxf(chi[k], p->c[k]);
      }
    }
    memcpy(n->lo, lo, sizeof lo);
    memcpy(n->hi, hi, sizeof hi);
    n->left = -1;
    if (n->count <= 4) continue;
    int best_axis = -1, best_bin = 0;
    float best = INFINITY;
    for (int ax = 0; ax < 3; ++ax) {
      float ext = chi[ax] - clo[ax];
      if (!(ext > 0.0f)) continue;
      float blo[NB][3], bhi[NB][3];
      uint32_t bc[NB] = {0};
      for (int b = 0; b < NB; ++b)
        for (int k = 0; k < 3; ++k) {
          blo[b][k] = INFINITY;
          bhi[b][k] = -INFINITY;
        }
      for (uint32_t i = n->first; i < n->first + n->count; ++i) {
        pref* p = &r[s->prim[i]];
        int b = (int)((p->c[ax] - clo[ax]) / ext * NB);
        if (b >= NB) b = NB - 1;
        bc[b]++;
        for (int k = 0; k < 3; ++k) {
          blo[b][k] = fminf(blo[b][k], p->lo[k]);
          bhi[b][k] = fmaxf(bhi[b][k], p->hi[k]);
        }
      }
      for (int sb = 0; sb < NB - 1; ++sb) {
        float l0[3] = {INFINITY, INFINITY, INFINITY}, l1[3] = {-INFINITY, -INFINITY, -INFINITY};
        float r0[3] = {INFINITY, INFINITY, INFINITY}, r1[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t nl = 0, nr = 0;
        for (int b = 0; b < NB; ++b) {
          if (!bc[b]) continue;
          float* x0 = b <= sb ? l0 : r0;
          float* x1 = b <= sb ? l1 : r1;
          for (int k = 0; k < 3; ++k) {
            x0[k] = fminf(x0[k], blo[b][k]);
            x1[k] = fmaxf(x1[k], bhi[b][k]);
          }
          if (b <= sb) nl += bc[b]; else nr += bc[b];
        }
        if (!nl || !nr) continue;
        float cost = box_area(l0, l1) * nl + box_area(r0, r1) * nr;
        if (cost < best) {
          best = cost;
          best_axis = ax;
          best_bin = sb;
        }
      }
    }
    uint32_t mid;
    if (best_axis < 0) {
      mid = n->first + n->count / 2;
    } else {
      float ext = chi[best_axis] - clo[best_axis];
      uint32_t i = n->first, j = n->first + n->count;
      while (i < j) {
        int b = (int)((r[s->prim[i]].c[best_axis] - clo[best_axis]) / ext * NB);
        if (b >= NB) b = NB - 1;
        if (b <= best_bin) {
          ++i;
        } else {
          uint32_t tmp = s->prim[i];
          s->prim[i] = s->prim[--j];
          s->prim[j] = tmp;
        }
      }
      mid = i;
      if (mid == n->first || mid == n->first + n->count) mid = n->first + n->count / 2;
    }
    uint32_t l = s->nn;
    s->nn += 2;
    bnode* L = &s->nodes[l];
    bnode* R = &s->nodes[l + 1];
    n = &s->nodes[ni];
    L->first = n->first;
    L->count = mid - n->first;
    R->first = mid;
    R->count = n->first + n->count - mid;
    n->left = (int32_t)l;
    st[sp++] = l + 1;
    st[sp++] = l;
  }
  free(st);
  free(r);
}

orc_scene* orc_scene_create(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* albedo,
                            const uint32_t* mat, const float* emission, const float* spec) {
  (void)nv;
  orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
  s->nt = nt;
  s->tri = (float*)malloc(sizeof(float) * 9 * nt);
  s->albedo = (float*)malloc(sizeof(float) * 4 * nt);
  memcpy(s->albedo, albedo, sizeof(float) * 4 * nt);
  if (spec) {
    s->spec = (float*)malloc(sizeof(float) * 4 * nt);
    memcpy(s->spec, spec, sizeof(float) * 4 * nt);
  }
  s->mat = (uint32_t*)malloc(sizeof(uint32_t) * nt);
  memcpy(s->mat, mat, sizeof(uint32_t) * nt);
  s->light_of = (int32_t*)malloc(sizeof(int32_t) * nt);
  for (uint32_t t = 0; t < nt; ++t) {
    const float* a = verts + 3 * idx[3 * t];
    const float* b = verts + 3 * idx[3 * t + 1];
    const float* c = verts + 3 * idx[3 * t + 2];
    float* q = s->tri + 9 * t;
    for (int k = 0; k < 3; ++k) {
      q[k] = a[k];
      q[3 + k] = b[k] - a[k];
      q[6 + k] = c[k] - a[k];
    }
    s->light_of[t] = -1;
  }
  /* emissive triangles with positive area, in triangle order, picked uniformly */
  uint32_t nl = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    const float* e = emission + 3 * t;
    if (e[0] + e[1] + e[2] > 0.0f) {
      const float* q = s->tri + 9 * t;
      float area = 0.5f * len(cross(ld3(q + 3), ld3(q + 6)));
      if (area > 0.0f) s->light_of[t] = (int32_t)nl++;
    }
  }
  s->nl = nl;
  s->light = (float*)malloc(sizeof(float) * 15 * (nl ? nl : 1));
  s->cdf = (float*)malloc(sizeof(float) * (nl ? nl : 1));
  s->light_tri = (uint32_t*)malloc(sizeof(uint32_t) * (nl ? nl : 1));
  double acc = 0.0;
  for (uint32_t t = 0; t < nt; ++t) {
    int32_t k = s->light_of[t];
    if (k < 0) continue;
    s->light_tri[k] = t;
    const float* q = s->tri + 9 * t;
    float* L = s->light + 15 * k;
    memcpy(L, q, 9 * sizeof(float));
    L[9] = 0.5f * len(cross(ld3(q + 3), ld3(q + 6)));
    L[10] = 1.0f / (float)nl;
    L[11] = emission[3 * t];
    L[12] = emission[3 * t + 1];
    L[13] = emission[3 * t + 2];
    acc += L[10];
    s->cdf[k] = (float)acc;
  }
  if (nl) s->cdf[nl - 1] = 1.0f;
  bvh_build(s);
  return s;
}

/* Light selection probabilities per light (light order = emissive triangles
 * of positive area in triangle order), e.g. the reference's uniform mesh then
 * uniform triangle choice (MeshLightList::Sample, SimpleKernel.h:114-120);
 * rebuilds the selection CDF the same way as the uniform default. */
void orc_scene_set_light_pdf(orc_scene* s, const float* pdf) {
  double acc = 0.0;
  for (uint32_t k = 0; k < s->nl; ++k) {
    s->light[15 * k + 10] = pdf[k];
    acc += pdf[k];
    s->cdf[k] = (float)acc;
  }
  if (s->nl) s->cdf[s->nl - 1] = 1.0f;
}

void orc_scene_destroy(orc_scene* s) {
  if (!s) return;
  free(s->tri);
  free(s->albedo);
  free(s->spec);
  free(s->mat);
  free(s->light_of);
  free(s->light);
  free(s->cdf);
  free(s->nodes);
  free(s->prim);
  free(s->light_tri);
  free(s->tc);
  free(s->tid);
  free(s->vn);
  free(s->texels);
  free(s->tex);
  free(s);
}

static void* dup_or_null(const void* src, size_t bytes) {
  if (!src) return NULL;
  void* d = malloc(bytes ? bytes : 1);
  memcpy(d, src, bytes);
  return d;
}

/* Texture coordinates, texture indices (0xFFFFFFFF = the constant white
 * texture), vertex normals and RGBA8 textures of the scene's triangles, as
 * the reference's hitgroup and mesh-light records carry them
 * (RTLibExtOPX7Test.h:1602-1638, RTLibExtOPX7TestApplication.cpp:262-280).
 * tex_id NULL: no texturing; vnormal NULL: glass shades with the face normal. */
void orc_scene_set_attributes(orc_scene* s, const float* texcoord, const uint32_t* tex_id, const float* vnormal,
                              const uint32_t* texels, uint64_t n_texels, const uint32_t* tex, uint32_t n_tex) {
  free(s->tc);
  free(s->tid);
  free(s->vn);
  free(s->texels);
  free(s->tex);
  s->tc = (float*)dup_or_null(tex_id ? texcoord : NULL, sizeof(float) * 6 * (size_t)s->nt);
  s->tid = (uint32_t*)dup_or_null(tex_id, sizeof(uint32_t) * 4 * (size_t)s->nt);
  s->vn = (float*)dup_or_null(vnormal, sizeof(float) * 9 * (size_t)s->nt);
  s->texels = (uint32_t*)dup_or_null(tex_id ? texels : NULL, sizeof(uint32_t) * (size_t)n_texels);
  s->tex = (uint32_t*)dup_or_null(tex_id ? tex : NULL, sizeof(uint32_t) * 4 * (size_t)n_tex);
}

/* Filtered fetch of a CUDA texture object with cudaFilterModeLinear,
 * normalised coordinates and cudaReadModeNormalizedFloat over RGBA8 texels,
 * the textures of TextureData (RTLibExtOPX7Test.h:396-437, 455-499), as the
 * CUDA C Programming Guide ("Texture Fetching", "Linear Filtering") states
 * it: wrap mode replaces u by frac(u), clamp mode clamps to [0, 1];
 * x = u W - 0.5, i = floor(x), alpha = frac(x) in 9-bit fixed point with 8
 * fractional bits (rounded to nearest even here); T[i], T[i + 1] with the
 * index wrapped or clamped; texel channels c / 255;
 * (1-a)(1-b) T00 + a(1-b) T10 + (1-a) b T01 + a b T11.  Rows are bottom-up
 * (LoadDefImage flips them, Core/Src/BinaryReader.cpp:768-782). */
void orc_tex_sample(const uint32_t* texels, const uint32_t* t, float u, float v, float* out3) {
  if (!isfinite(u)) u = 0.0f;
  if (!isfinite(v)) v = 0.0f;
  if (t[3]) {
    u = u - floorf(u);
    v = v - floorf(v);
  } else {
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
  }
  const int w = (int)t[1], h = (int)t[2];
  float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = rintf((x - fx) * 256.0f) * (1.0f / 256.0f);
  float b = rintf((y - fy) * 256.0f) * (1.0f / 256.0f);
  int i[2] = {(int)fx, (int)fx + 1}, j[2] = {(int)fy, (int)fy + 1};
  for (int k = 0; k < 2; ++k) {
    if (t[3]) {
      i[k] = (i[k] % w + w) % w;
      j[k] = (j[k] % h + h) % h;
    } else {
      i[k] = i[k] < 0 ? 0 : (i[k] >= w ? w - 1 : i[k]);
      j[k] = j[k] < 0 ? 0 : (j[k] >= h ? h - 1 : j[k]);
    }
  }
  const uint32_t* base = texels + t[0];
  uint32_t c00 = base[(size_t)j[0] * w + i[0]], c10 = base[(size_t)j[0] * w + i[1]];
  uint32_t c01 = base[(size_t)j[1] * w + i[0]], c11 = base[(size_t)j[1] * w + i[1]];
  float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
  for (int k = 0; k < 3; ++k) {
    int sh = 8 * k;
    float t00 = (float)((c00 >> sh) & 0xFFu) / 255.0f, t10 = (float)((c10 >> sh) & 0xFFu) / 255.0f;
    float t01 = (float)((c01 >> sh) & 0xFFu) / 255.0f, t11 = (float)((c11 >> sh) & 0xFFu) / 255.0f;
    out3[k] = w00 * t00 + w10 * t10 + w01 * t01 + w11 * t11;
  }
}

/* Texture coordinate at barycentrics (b0, b1, b2) (GetTexCrd,
 * SimpleKernel.h:70-77, 251-257) and the factor of texture slot `slot` of
 * triangle `tri` there (1 without a texture). */
static v3 tex_factor(const orc_scene* s, int64_t tri, int slot, float b0, float b1, float b2) {
  if (!s->tid) return V(1, 1, 1);
  uint32_t id = s->tid[4 * tri + slot];
  if (id == 0xFFFFFFFFu) return V(1, 1, 1);
  const float* c = s->tc + 6 * tri;
  float u = b0 * c[0] + b1 * c[2] + b2 * c[4];
  float v = b0 * c[1] + b1 * c[3] + b2 * c[5];
  float o[3];
  orc_tex_sample(s->texels, s->tex + 4 * (size_t)id, u, v, o);
  return V(o[0], o[1], o[2]);
}

static int box_hit(const bnode* n, v3 o, v3 inv, float tmin, float tmax) {
  float t0 = tmin, t1 = tmax;
  float ov[3] = {o.x, o.y, o.z}, iv[3] = {inv.x, inv.y, inv.z};
  for (int k = 0; k < 3; ++k) {
    float a = (n->lo[k] - ov[k]) * iv[k], b = (n->hi[k] - ov[k]) * iv[k];
    if (a > b) {
      float tt = a;
      a = b;
      b = tt;
    }
    t0 = fmaxf(t0, a);
    t1 = fminf(t1, b * 1.00000024f);
  }
  return t0 <= t1;
}

/* closest (any = 0) or any hit (any = 1); returns triangle or -1 */
static int64_t trace_uv(const orc_scene* s, v3 o, v3 d, float tmin, float tmax, int any, float* t_out, uint64_t* rays,
                        float* u_out, float* v_out) {
  (*rays)++;
  const float eps = 1e-20f;
  v3 dd = V(fabsf(d.x) > eps ? d.x : copysignf(eps, d.x), fabsf(d.y) > eps ? d.y : copysignf(eps, d.y),
            fabsf(d.z) > eps ? d.z : copysignf(eps, d.z));
  v3 inv = V(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
  float tb = tmax, u, v;
  int64_t best = -1;
  uint32_t st[128];
  int sp = 0;
  st[sp++] = 0;
  float ov[3] = {o.x, o.y, o.z}, dv[3] = {d.x, d.y, d.z};
  while (sp) {
    const bnode* n = &s->nodes[st[--sp]];
    if (!box_hit(n, o, inv, tmin, tb)) continue;
    if (n->left < 0) {
      for (uint32_t i = n->first; i < n->first + n->count; ++i) {
        uint32_t t = s->prim[i];
        const float* q = s->tri + 9 * t;
        if (orc_isect_tri(ov, dv, q, q + 3, q + 6, tmin, &tb, &u, &v)) {
          best = t;
          if (u_out) {
            *u_out = u;
            *v_out = v;
          }
          if (any) {
            *t_out = tb;
            return best;
          }
        }
      }
    } else if (sp + 2 <= 128) {
      st[sp++] = (uint32_t)n->left + 1;
      st[sp++] = (uint32_t)n->left;
    }
  }
  *t_out = tb;
  return best;
}

static int64_t trace(const orc_scene* s, v3 o, v3 d, float tmin, float tmax, int any, float* t_out, uint64_t* rays) {
  return trace_uv(s, o, d, tmin, tmax, any, t_out, rays, NULL, NULL);
}

/* ======================================================================
 * integrator — the same estimator, random-number order and guide logic as the
 * wavefront kernels (NEE with one light sample or RIS, MIS by the power
 * heuristic, Russian roulette, SD-tree guided mixture and record).
 * ==================================================================== */
enum { F_NEE = 1, F_RIS = 2, F_USE_GRID = 4, F_USE_TREE = 8, F_GUIDE = 16, F_FINAL = 32, F_LOCATE = 64, F_MIS = 128,
       F_RR = 256 };
enum { M_LAMBERT = 0, M_GGX = 1, M_PHONG = 2, M_GLASS = 3 };

static float ggx_d(float c, float a2) {
  float c2 = c * c;
  float den = c2 * (a2 - 1.0f) + 1.0f;
  return a2 / (PI_F * den * den);
}
static float ggx_g1(float cv, float a2) {
  float c = fabsf(cv);
  float c2 = c * c;
  return 2.0f * c / (c + sqrtf(a2 + (1.0f - a2) * c2));
}
static v3 schlick3(v3 f0, float c) {
  float m = 1.0f - fminf(fmaxf(c, 0.0f), 1.0f);
  float m5 = m * m * m * m * m;
  return add(f0, scl(sub(V(1, 1, 1), f0), m5));
}
static float phong_lobe(v3 dir, v3 refl, float shin) {
  float rc = fmaxf(dot(refl, dir), 0.0f);
  float p = (shin != 0.0f) ? powf(rc, shin) : (rc == 0.0f ? 1.0f : 0.0f);
  return (shin + 1.0f) * p * INV_2PI_F;
}

typedef struct {
  v3 f;
  float pdf;
} beval;

static beval bsdf_eval(uint32_t type, v3 albedo, float param, v3 spec, v3 n, v3 wo, v3 wi) {
  beval e;
  float ci = dot(n, wi);
  if (type == M_GGX) {
    float co = dot(n, wo);
    if (ci <= 0.0f || co <= 0.0f) {
      e.f = V(0, 0, 0);
      e.pdf = 0.0f;
      return e;
    }
    float a = fmaxf(param, 1e-3f), a2 = a * a;
    v3 h = nrm(add(wi, wo));
    float ch = dot(n, h);
    float D = ggx_d(ch, a2);
    float G = ggx_g1(ci, a2) * ggx_g1(co, a2);
    float oh = fmaxf(dot(wo, h), 1e-8f);
    v3 fr = schlick3(albedo, oh);
    e.f = scl(fr, D * G / (4.0f * ci * co));
    e.pdf = D * ch / (4.0f * oh);
    return e;
  }
  if (type == M_PHONG) {
    v3 refl = nrm(add(scl(wo, -1.0f), scl(n, 2.0f * dot(n, wo))));
    float ad = 1.0e-10f + avg3(albedo), as = 1.0e-10f + avg3(spec);
    float sel = ad / (ad + as);
    float lobe = phong_lobe(wi, refl, param);
    e.f = add(scl(albedo, INV_PI_F), scl(spec, ((param + 2.0f) / (param + 1.0f)) * lobe));
    e.pdf = sel * fmaxf(ci * INV_PI_F, 0.0f) + (1.0f - sel) * lobe;
    if (ci <= 0.0f) e.f = V(0, 0, 0);
    return e;
  }
  if (ci <= 0.0f) {
    e.f = V(0, 0, 0);
    e.pdf = 0.0f;
    return e;
  }
  e.f = scl(albedo, INV_PI_F);
  e.pdf = ci * INV_PI_F;
  return e;
}

static v3 cosine_local(float u1, float u2) {
  float r = sqrtf(u1), z = sqrtf(1.0f - u1), t = 2.0f * u2 * PI_F;
  return V(r * cosf(t), r * sinf(t), z);
}

static v3 bsdf_sample(uint32_t type, v3 albedo, float param, v3 spec, v3 n, v3 wo, pcg32* r) {
  if (type == M_GGX) {
    float u1 = rnd(r), u2 = rnd(r);
    float a = fmaxf(param, 1e-3f), a2 = a * a;
    float ch = sqrtf((1.0f - u1) / (1.0f + (a2 - 1.0f) * u1));
    float sh = sqrtf(fmaxf(1.0f - ch * ch, 0.0f));
    float phi = TWO_PI_F * u2;
    onb o = make_onb(n);
    v3 h = onb_local(&o, V(sh * cosf(phi), sh * sinf(phi), ch));
    return sub(scl(h, 2.0f * dot(wo, h)), wo);
  }
  if (type == M_PHONG) {
    float us = rnd(r), u1 = rnd(r), u2 = rnd(r);
    float ad = 1.0e-10f + avg3(albedo), as = 1.0e-10f + avg3(spec);
    float sel = ad / (ad + as);
    if (us < sel) {
      onb o = make_onb(n);
      return onb_local(&o, cosine_local(u1, u2));
    }
    v3 refl = nrm(add(scl(wo, -1.0f), scl(n, 2.0f * dot(n, wo))));
    float ct = powf(u1, 1.0f / (param + 1.0f));
    float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
    float phi = TWO_PI_F * u2;
    onb o = make_onb(refl);
    return onb_local(&o, V(st * cosf(phi), st * sinf(phi), ct));
  }
  float u1 = rnd(r), u2 = rnd(r);
  onb o = make_onb(n);
  return onb_local(&o, cosine_local(u1, u2));
}

static float power_h(float a, float b) {
  float a2 = a * a, b2 = b * b;
  return (a2 + b2) > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

static uint32_t pick_light(const orc_scene* s, float u) {
  uint32_t lo = 0, hi = s->nl - 1u;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (s->cdf[mid] > u) hi = mid; else lo = mid + 1u;
  }
  return lo;
}

typedef struct {
  v3 p, n, le;
  float area, pdf;
  uint32_t idx;
} lsample;

/* Per-hit terms of one path for the adjoint: emission and NEE contributions
 * with their lights, and the Lambertian triangle that scattered at each hit. */
typedef struct {
  uint32_t nh;
  v3 e[64];
  int32_t el[64];
  v3 n[64];
  int32_t nl[64];
  int32_t tri[64];
  v3 rho[64];
} adjrec;

static lsample sample_light(const orc_scene* s, pcg32* r) {
  float ul = rnd(r), u1 = rnd(r), u2 = rnd(r);
  uint32_t li = pick_light(s, ul);
  const float* L = s->light + 15 * li;
  float lo = fminf(u1, u2), hi = fmaxf(u1, u2);
  lsample ls;
  ls.idx = li;
  v3 v0 = ld3(L), e1 = ld3(L + 3), e2 = ld3(L + 6);
  ls.p = add(add(v0, scl(e1, 1.0f - hi)), scl(e2, hi - lo));
  ls.n = nrm(cross(e1, e2));
  ls.le = ld3(L + 11);
  /* MeshLight::GetEmission at the sampled point (SimpleKernel.h:81-104):
   * barycentrics (lo, 1 - hi, hi - lo) */
  if (s->tid) ls.le = mul(ls.le, tex_factor(s, s->light_tri[li], 3, lo, 1.0f - hi, hi - lo));
  ls.area = L[9];
  ls.pdf = L[10];
  return ls;
}

typedef struct {
  v3 T;       /* throughput after the bounce */
  v3 pos;     /* hit point (for the spatial box filter) */
  float wo_pdf, cosv, cu, cv;
  uint32_t cell, flags;
} vtx;

enum { VTX_VALID = 1, VTX_DELTA = 2, VTX_BAD_BSDF = 4 };

static uint64_t render_path(const orc_scene* s, const orc_params* p, const orc_sdtree* g, const orc_hashguide* hg,
                            uint32_t pix, uint32_t smp, v3* Lout, adjrec* ar) {
  uint64_t rays = 0;
  pcg32 r = path_rng(p->seed, pix, p->sample_offset + smp);
  uint32_t x = pix % p->width, y = pix / p->width;
  float jx = rnd(&r), jy = rnd(&r);
  float dx = 2.0f * ((float)x + jx) / (float)p->width - 1.0f;
  float dy = 2.0f * ((float)y + jy) / (float)p->height - 1.0f;
  v3 o = ld3(p->eye);
  v3 d = nrm(add(add(scl(ld3(p->u), dx), scl(ld3(p->v), dy)), ld3(p->w)));
  v3 T = V(1, 1, 1), L = V(0, 0, 0);
  float prev_pdf = 0.0f;
  int prev_delta = 1;
  const int nee = (p->flags & F_NEE) && s->nl > 0;
  const int ris = nee && (p->flags & F_RIS);
  const int mis = nee && !ris && (p->flags & F_MIS);
  const int use_tree = g != NULL && (p->flags & F_USE_TREE);
  /* the hash-grid guide of SimpleGuide2.cu (USE_GRID) */
  const int use_grid = !use_tree && hg != NULL && (p->flags & F_USE_GRID);
  const int record = (use_tree || use_grid) && !(p->flags & F_FINAL);
  const float alpha = p->fraction;
  v3 cbuf[64];
  vtx vb[64];
  uint32_t nh = 0;
  for (uint32_t depth = 0; depth < p->max_depth; ++depth) {
    float t, bu = 0.0f, bv = 0.0f;
    int64_t tri = trace_uv(s, o, d, 1e-4f, 1e20f, 0, &t, &rays, &bu, &bv);
    v3 c = V(0, 0, 0);
    vtx vx;
    memset(&vx, 0, sizeof vx);
    vx.cell = INVALID;
    int cont = 0;
    v3 Tn = V(0, 0, 0), nd = V(0, 0, 0), pos = V(0, 0, 0);
    float wo_pdf = 0.0f;
    int next_delta = 0;
    if (ar) {
      ar->e[nh] = V(0, 0, 0);
      ar->n[nh] = V(0, 0, 0);
      ar->el[nh] = ar->nl[nh] = ar->tri[nh] = -1;
    }
    if (tri < 0) {
      c = mul(T, ld3(p->bg));
      if (ar) ar->e[nh] = c;  /* background: no light parameter, but depends on earlier albedos */
    } else {
      const float* q = s->tri + 9 * tri;
      v3 ng = nrm(cross(ld3(q + 3), ld3(q + 6)));
      float dn = dot(d, ng);
      int front = dn < 0.0f;
      v3 n = front ? ng : neg(ng);
      pos = add(o, scl(d, t));
      v3 wo = neg(d);
      if (s->light_of[tri] >= 0) {
        if (front) {
          const float* Lr = s->light + 15 * s->light_of[tri];
          float w = 1.0f;
          if (depth > 0 && !prev_delta && nee) {
            if (mis) {
              float pl = Lr[10] / Lr[9] * t * t / fabsf(dn);
              w = power_h(prev_pdf, pl);
            } else {
              w = 0.0f;
            }
          }
          /* SampleEmission at the hit's texture coordinate */
          v3 le = mul(ld3(Lr + 11), tex_factor(s, tri, 2, 1.0f - bu - bv, bu, bv));
          c = scl(mul(T, le), w);
          if (ar) {
            ar->e[nh] = c;
            ar->el[nh] = s->light_of[tri];
          }
        }
      } else if (depth + 1u < p->max_depth) {
        uint32_t type = s->mat[tri];
        v3 albedo = ld3(s->albedo + 4 * tri);
        float param = s->albedo[4 * tri + 3];
        if (ar && type == M_LAMBERT) {
          ar->tri[nh] = (int32_t)tri;
          ar->rho[nh] = albedo;
        }
        v3 spec = V(0, 0, 0);
        if (type == M_PHONG && s->spec) {
          spec = ld3(s->spec + 4 * tri);
          param = s->spec[4 * tri + 3];
        }
        /* SampleDiffuse / SampleSpecular (SimpleKernel.h:251-272); glass keeps
         * its specular colour in albedo */
        albedo = mul(albedo, tex_factor(s, tri, type == M_GLASS ? 1 : 0, 1.0f - bu - bv, bu, bv));
        if (type == M_PHONG) spec = mul(spec, tex_factor(s, tri, 1, 1.0f - bu - bv, bu, bv));
        if (type == M_GLASS) {
          /* the interpolated vertex normal when the scene has them
           * (GetTriangleVNormal, SimpleKernel.h:238-250) */
          if (s->vn) {
            const float* q9 = s->vn + 9 * tri;
            float b0 = 1.0f - bu - bv;
            ng = nrm(add(add(scl(ld3(q9), b0), scl(ld3(q9 + 3), bu)), scl(ld3(q9 + 6), bv)));
          }
          float ior = param, ci = dot(ng, d), eta;
          v3 rn;
          if (ci < 0.0f) {
            rn = ng;
            eta = 1.0f / ior;
            ci = -ci;
          } else {
            rn = neg(ng);
            eta = ior;
          }
          float so2 = (1.0f - ci * ci) * eta * eta;
          float f0 = (1.0f - eta) / (1.0f + eta);
          f0 = f0 * f0;
          float m = 1.0f - ci;
          float fres = f0 + (1.0f - f0) * (m * m * m * m * m);
          v3 refl = nrm(sub(d, scl(rn, 2.0f * dot(d, rn))));
          if (rnd(&r) < fres || so2 > 1.0f) {
            pos = add(pos, scl(rn, 1e-4f));
            nd = refl;
            Tn = mul(T, albedo);
          } else {
            pos = sub(pos, scl(rn, 1e-4f));
            float si2 = fmaxf(1.0f - ci * ci, 0.0f);
            float co = sqrtf(1.0f - so2);
            if (si2 > 0.0f) {
              v3 k = divs(add(d, scl(rn, ci)), sqrtf(si2));
              nd = nrm(sub(scl(k, sqrtf(so2)), scl(rn, co)));
            } else {
              nd = d;
            }
            Tn = T;
          }
          wo_pdf = fabsf(dot(nd, rn));
          vx.cosv = wo_pdf;
          vx.flags = VTX_DELTA;
          next_delta = 1;
          cont = max3(Tn) > 0.0f;
        } else {
          /* cell: where this vertex records; gcell: the guide it samples from.
           * The hash grid (SimpleGuide2.cu:319-323, 420-428) samples the
           * previous table's slot and records there, or, in a Locate pass,
           * into the slot a position missing from the previous table inserts
           * into the current one. */
          uint32_t cell = INVALID, gcell = INVALID;
          if (use_tree) {
            cell = stree_lookup1(g->stree, g->aabb_min, g->size, pos);
            gcell = cell;
          } else if (use_grid) {
            float pp[3] = {pos.x, pos.y, pos.z};
            uint32_t c3[3], bk, cs;
            orc_grid_cell_hash(&hg->grid, pp, c3, &bk, &cs);
            gcell = grid_find_prv1(hg->prv, bk, cs);
            cell = gcell;
            if ((p->flags & F_LOCATE) && gcell == INVALID) cell = grid_find_cur1(hg->cur, bk, cs);
          }
          const int guided = (use_tree || use_grid) && (p->flags & F_GUIDE) && gcell != INVALID;
          const float* gs = guided && use_tree ? g->s_sums + 4 * (size_t)g->s_base[cell] : NULL;
          const uint16_t* gc = guided && use_tree ? g->s_child + 4 * (size_t)g->s_base[cell] : NULL;
          const float* mw = guided && use_grid ? hg->s_w + (size_t)hg->per_cell * gcell : NULL;
          if (nee) {
            if (ris) {
              float wsum = 0.0f, fa_y = 0.0f, dist_y = 0.0f;
              uint32_t m = 0;
              v3 f_y = V(0, 0, 0), dir_y = V(0, 0, 0);
              for (uint32_t k = 0; k < p->num_candidates; ++k) {
                lsample ls = sample_light(s, &r);
                v3 dl = sub(ls.p, pos);
                float dist = len(dl);
                v3 dir = divs(dl, dist);
                float ndl = dot(dir, n), lndl = -dot(dir, ls.n);
                beval be = bsdf_eval(type, albedo, param, spec, n, wo, dir);
                float gg = fmaxf(ndl, 0.0f) * fmaxf(lndl, 0.0f) / (dist * dist);
                v3 f = scl(mul(be.f, ls.le), gg);
                float fa = avg3(f);
                float wi = fa * (ls.area / ls.pdf);
                wsum += wi;
                ++m;
                float ur = rnd(&r);
                if ((wi / wsum) >= ur) {
                  f_y = f;
                  fa_y = fa;
                  dir_y = dir;
                  dist_y = dist;
                }
              }
              if (wsum > 0.0f && fa_y > 0.0f) {
                float W = wsum / (fa_y * (float)m);
                v3 con = scl(mul(T, f_y), W);
                float tt;
                if (max3(con) > 0.0f && trace(s, pos, dir_y, 1e-4f, dist_y - 1e-4f, 1, &tt, &rays) < 0) c = add(c, con);
              }
            } else {
              lsample ls = sample_light(s, &r);
              v3 dl = sub(ls.p, pos);
              float dist2 = dot(dl, dl), dist = sqrtf(dist2);
              v3 dir = divs(dl, dist);
              float cl = -dot(dir, ls.n), cs = dot(dir, n);
              if (cl > 0.0f && cs > 0.0f) {
                beval be = bsdf_eval(type, albedo, param, spec, n, wo, dir);
                float pl = ls.pdf / ls.area * dist2 / cl;
                float w = 1.0f;
                if (mis) {
                  float bp = be.pdf;
                  if (guided) {
                    float cu, cv;
                    dir_to_canonical(dir, &cu, &cv);
                    float gp = use_tree ? dtree_pdf1(gs, gc, g->s_sum[cell], g->s_sw[cell], cu, cv)
                                        : mq_pdf1(mw, hg->level, cu, cv);
                    bp = (1.0f - alpha) * bp + alpha * gp;
                  }
                  w = power_h(pl, bp);
                }
                v3 con = scl(mul(mul(T, be.f), ls.le), cs * w / pl);
                float tt;
                if (max3(con) > 0.0f && trace(s, pos, dir, 1e-4f, dist - 1e-4f, 1, &tt, &rays) < 0) {
                  c = add(c, con);
                  if (ar) {
                    ar->n[nh] = con;
                    ar->nl[nh] = (int32_t)ls.idx;
                  }
                }
              }
            }
          }
          v3 dir;
          float tree_pdf = -1.0f; /* density of a direction drawn from the tree */
          if (guided) {
            float us = rnd(&r);
            if (us < alpha) {
              float gu, gv;
              if (use_tree) {
                dtree_sample1(gs, gc, g->s_sum[cell], g->s_sw[cell], &r, &gu, &gv, &tree_pdf);
              } else {
                /* MortonQuadTreeWrapperT::SampleAndPdf (:486-500): the density
                 * of the sampled bin */
                tree_pdf = mq_sample1(mw, hg->level, &r, &gu, &gv);
              }
              dir = nrm(canonical_to_dir(gu, gv));
            } else {
              dir = bsdf_sample(type, albedo, param, spec, n, wo, &r);
            }
          } else {
            dir = bsdf_sample(type, albedo, param, spec, n, wo, &r);
          }
          beval be = bsdf_eval(type, albedo, param, spec, n, wo, dir);
          wo_pdf = be.pdf;
          if (guided) {
            dir_to_canonical(dir, &vx.cu, &vx.cv);
            float gp = tree_pdf >= 0.0f ? tree_pdf
                       : use_tree ? dtree_pdf1(gs, gc, g->s_sum[cell], g->s_sw[cell], vx.cu, vx.cv)
                                  : mq_pdf1(mw, hg->level, vx.cu, vx.cv);
            wo_pdf = (1.0f - alpha) * be.pdf + alpha * fmaxf(gp, 0.0f);
          } else if (record && cell != INVALID) {
            dir_to_canonical(dir, &vx.cu, &vx.cv);
          }
          float cd = dot(dir, n);
          vx.cosv = cd;
          Tn = (wo_pdf > 0.0f && cd > 0.0f) ? scl(mul(T, be.f), cd / wo_pdf) : V(0, 0, 0);
          if ((p->flags & F_RR) && depth >= p->rr_depth) {
            float qq = fminf(max3(Tn), 0.95f);
            if (rnd(&r) >= qq) {
              Tn = V(0, 0, 0);
            } else {
              Tn = divs(Tn, qq);
            }
          }
          int ok = finite3(be.f) && be.f.x >= 0.0f && be.f.y >= 0.0f && be.f.z >= 0.0f;
          vx.flags = ok ? 0u : VTX_BAD_BSDF;
          vx.cell = cell;
          vx.pos = pos;
          nd = dir;
          cont = max3(Tn) > 0.0f && isfinite(max3(Tn));
        }
      }
    }
    L = add(L, c);
    cbuf[nh] = c;
    vx.T = Tn;
    vx.wo_pdf = wo_pdf;
    if (cont && vx.cell != INVALID) vx.flags |= VTX_VALID;
    vb[nh] = vx;
    nh++;
    if (!cont) break;
    o = pos;
    d = nd;
    T = Tn;
    prev_pdf = wo_pdf;
    prev_delta = next_delta;
  }
  if (record) {
    /* TraceVertexT::Commit :830-866 over vertices 0 .. nh-2 */
    for (uint32_t dv = 0; dv + 1 < nh; ++dv) {
      vtx* v = &vb[dv];
      if (!(v->flags & VTX_VALID) || (v->flags & VTX_DELTA) || v->cell == INVALID) continue;
      float rx = 0.0f, ry = 0.0f, rz = 0.0f;
      for (uint32_t k = dv + 1; k < nh; ++k) {
        rx += cbuf[k].x;
        ry += cbuf[k].y;
        rz += cbuf[k].z;
      }
      int ok = isfinite(rx) && rx >= 0.0f && isfinite(ry) && ry >= 0.0f && isfinite(rz) && rz >= 0.0f;
      if (v->wo_pdf <= 0.0f || !ok || (v->flags & VTX_BAD_BSDF)) continue;
      float lx = v->T.x * v->wo_pdf > 1e-4f ? rx / v->T.x : 0.0f;
      float ly = v->T.y * v->wo_pdf > 1e-4f ? ry / v->T.y : 0.0f;
      float lz = v->T.z * v->wo_pdf > 1e-4f ? rz / v->T.z : 0.0f;
      float cs = fabsf(v->cosv);
      lx *= cs;
      ly *= cs;
      lz *= cs;
      float value = ((lx + ly + lz) / 3.0f) / v->wo_pdf;
      uint32_t cell = v->cell;
      float cu = v->cu, cv = v->cv;
      if (use_grid) {
        /* MortonTraceVertexT::Commit (OPX7UtilsMorton.h:550-585) ->
         * MortonQuadTreeT::Record, leaf-only branch (:262-285) */
        uint32_t nn = 1u << hg->level;
        size_t at = (size_t)hg->per_cell * cell + level_offset(hg->level) + orc_morton2(canon_bin(cu, nn), canon_bin(cv, nn));
        atomic_addf(&hg->b_w[at], value);
        continue;
      }
      if (g->filter) {
        /* stochastic box filters, own stream per (pixel, sample, vertex depth) */
        pcg32 fr = filter_rng(p->seed, pix, p->sample_offset + smp, dv);
        float f[5];
        for (int k = 0; k < 5; ++k) f[k] = rnd(&fr);
        if (g->filter & FILTER_SPATIAL) cell = stree_box_jitter1(g->stree, g->aabb_min, g->size, v->pos, f[0], f[1], f[2]);
        if (cell == INVALID) continue;
        if ((g->filter & FILTER_DIRECTIONAL) && !dtree_box_jitter1(g->b_child, g->b_base[cell], &cu, &cv, f[3], f[4]))
          value = 0.0f;
      }
      dtree_record1(g->b_sw, g->b_sums, g->b_child, g->b_base[cell], cell, cu, cv, value, 1.0f);
    }
  }
  *Lout = L;
  if (ar) ar->nh = nh;
  return rays;
}

/* Reference gradient of the path-replay adjoint, accumulated in double:
 * d/d rho_t = sum over Lambertian visits v of t of (NEE at v + all
 * contributions after v) / rho; d/d Le_l = contributions of light l / Le_l.
 * Weighted by dl_dimage / spp per pixel.  RIS and Russian roulette are not
 * differentiated (as in the GPU adjoint). */
void orc_adjoint(const orc_scene* s, const orc_params* p, const float* dldi, double* grad_alb, double* grad_le) {
  adjrec* ar = (adjrec*)malloc(sizeof(adjrec));
  for (uint32_t pix = p->px_begin; pix < p->px_end; ++pix) {
    double w[3] = {dldi[3 * pix] / (double)p->spp, dldi[3 * pix + 1] / (double)p->spp,
                   dldi[3 * pix + 2] / (double)p->spp};
    for (uint32_t smp = 0; smp < p->spp; ++smp) {
      v3 L;
      render_path(s, p, NULL, NULL, pix, smp, &L, ar);
      for (uint32_t k = 0; k < ar->nh; ++k) {
        const float* cs[2] = {&ar->e[k].x, &ar->n[k].x};
        int32_t ls[2] = {ar->el[k], ar->nl[k]};
        for (int q = 0; q < 2; ++q) {
          if (ls[q] < 0) continue;
          const float* le = s->light + 15 * ls[q] + 11;
          for (int ch = 0; ch < 3; ++ch)
            if (le[ch] > 0.0f) grad_le[3 * ls[q] + ch] += w[ch] * (double)cs[q][ch] / (double)le[ch];
        }
        if (ar->tri[k] < 0) continue;
        double aff[3];
        aff[0] = ar->n[k].x;
        aff[1] = ar->n[k].y;
        aff[2] = ar->n[k].z;
        for (uint32_t j = k + 1; j < ar->nh; ++j) {
          aff[0] += (double)ar->e[j].x + (double)ar->n[j].x;
          aff[1] += (double)ar->e[j].y + (double)ar->n[j].y;
          aff[2] += (double)ar->e[j].z + (double)ar->n[j].z;
        }
        const float* rho = &ar->rho[k].x;
        for (int ch = 0; ch < 3; ++ch)
          if (rho[ch] > 0.0f) grad_alb[3 * ar->tri[k] + ch] += w[ch] * aff[ch] / (double)rho[ch];
      }
    }
  }
  free(ar);
}

typedef struct {
  const orc_scene* s;
  const orc_params* p;
  const orc_sdtree* g;
  const orc_hashguide* hg;
  float* accum;
  uint32_t next;
  uint64_t rays;
  pthread_mutex_t mu;
} job;

static void* worker(void* arg) {
  job* j = (job*)arg;
  uint64_t rays = 0;
  const uint32_t tile = 256;
  for (;;) {
    uint32_t b = __atomic_fetch_add(&j->next, tile, __ATOMIC_RELAXED);
    uint32_t begin = j->p->px_begin + b;
    if (begin >= j->p->px_end) break;
    uint32_t end = begin + tile < j->p->px_end ? begin + tile : j->p->px_end;
    for (uint32_t pix = begin; pix < end; ++pix) {
      float sx = 0.0f, sy = 0.0f, sz = 0.0f;
      for (uint32_t smp = 0; smp < j->p->spp; ++smp) {
        v3 L;
        rays += render_path(j->s, j->p, j->g, j->hg, pix, smp, &L, NULL);
        sx += L.x;
        sy += L.y;
        sz += L.z;
      }
      j->accum[3 * pix] += sx;
      j->accum[3 * pix + 1] += sy;
      j->accum[3 * pix + 2] += sz;
    }
  }
  pthread_mutex_lock(&j->mu);
  j->rays += rays;
  pthread_mutex_unlock(&j->mu);
  return NULL;
}

/* Debug frames: __raygen__debug / __closesthit__debug / __miss__debug
 * (SimpleTrace.cu:104-213, 631-734) for DebugFrameType (SimpleKernel.h:160-171)
 * 0 normal, 1 distance, 2 diffuse, 3 specular, 4 emission, 5 shininess, 6
 * refraction index, and 8 world position (7, the grid value, is left to the
 * caller): per pixel the raw quantity and the hit triangle (-1 on a miss) of
 * the camera ray of sample p->sample_offset, traced in [1e-4, 1e20].  A miss
 * gives normal (1,0,0), distance 1e20, zeros otherwise. */
void orc_debug_frame(const orc_scene* s, const orc_params* p, uint32_t type, float* out3, int32_t* tri_out) {
  uint32_t npix = p->width * p->height;
  for (uint32_t pix = 0; pix < npix; ++pix) {
    pcg32 r = path_rng(p->seed, pix, p->sample_offset);
    uint32_t x = pix % p->width, y = pix / p->width;
    float jx = rnd(&r), jy = rnd(&r);
    float dx = 2.0f * ((float)x + jx) / (float)p->width - 1.0f;
    float dy = 2.0f * ((float)y + jy) / (float)p->height - 1.0f;
    v3 o = ld3(p->eye);
    v3 d = nrm(add(add(scl(ld3(p->u), dx), scl(ld3(p->v), dy)), ld3(p->w)));
    float t = 0.0f, bu = 0.0f, bv = 0.0f;
    uint64_t rays = 0;
    int64_t tri = trace_uv(s, o, d, 1e-4f, 1e20f, 0, &t, &rays, &bu, &bv);
    v3 q = V(0, 0, 0);
    if (tri < 0) {
      if (type == 0) q = V(1, 0, 0);
      if (type == 1) q = V(1e20f, 1e20f, 1e20f);
    } else {
      const float* tq = s->tri + 9 * tri;
      const float* al = s->albedo + 4 * tri;
      const float* sp = s->spec ? s->spec + 4 * tri : NULL;
      uint32_t mt = s->mat[tri];
      int light = s->light_of[tri] >= 0;
      if (type == 0) q = nrm(cross(ld3(tq + 3), ld3(tq + 6)));
      if (type == 1) q = V(t, t, t);
      float b0 = 1.0f - bu - bv;
      if (type == 2 && mt != M_GLASS && !light) q = mul(ld3(al), tex_factor(s, tri, 0, b0, bu, bv));
      if (type == 3 && mt == M_GLASS) q = mul(ld3(al), tex_factor(s, tri, 1, b0, bu, bv));
      if (type == 3 && mt == M_PHONG && sp) q = mul(ld3(sp), tex_factor(s, tri, 1, b0, bu, bv));
      if (type == 4 && light) q = mul(ld3(s->light + 15 * s->light_of[tri] + 11), tex_factor(s, tri, 2, b0, bu, bv));
      if (type == 5 && mt == M_PHONG && sp) q = V(sp[3], sp[3], sp[3]);
      if (type == 6 && mt == M_GLASS) q = V(al[3], al[3], al[3]);
      if (type == 8) q = add(o, scl(d, t));
    }
    out3[3 * pix] = q.x;
    out3[3 * pix + 1] = q.y;
    out3[3 * pix + 2] = q.z;
    tri_out[pix] = (int32_t)tri;
  }
}

static uint64_t render_job(const orc_scene* s, const orc_params* p, const orc_sdtree* guide,
                           const orc_hashguide* hg, float* accum) {
  job j;
  j.s = s;
  j.p = p;
  j.g = guide;
  j.hg = hg;
  j.accum = accum;
  j.next = 0;
  j.rays = 0;
  pthread_mutex_init(&j.mu, NULL);
  uint32_t nth = p->threads ? p->threads : 1;
  if (nth > 256) nth = 256;
  pthread_t th[256];
  for (uint32_t i = 0; i < nth; ++i) pthread_create(&th[i], NULL, worker, &j);
  for (uint32_t i = 0; i < nth; ++i) pthread_join(th[i], NULL);
  pthread_mutex_destroy(&j.mu);
  return j.rays;
}

uint64_t orc_render(const orc_scene* s, const orc_params* p, const orc_sdtree* guide, float* accum) {
  return render_job(s, p, guide, NULL, accum);
}

uint64_t orc_render_grid(const orc_scene* s, const orc_params* p, const orc_hashguide* guide, float* accum) {
  return render_job(s, p, NULL, guide, accum);
}

/* ======================================================================
 * pass controllers: RTSTreeControllerT::BegTrace/EndTrace
 * (Utils/OPX7UtilsPathGuiding.h:1805-1882) and RTMortonQuadTreeControllerT::
 * BegTrace/EndTrace (Utils/OPX7UtilsMorton.h:806-905), one launch at a time
 * as RTLibExtOPX7TestApplication::TracePipeline drives them (:2155-2231,
 * including its hash-grid snapshot when GetSamplePerTmp() is 0 after EndTrace).
 * ==================================================================== */
uint32_t orc_ctrl_schedule(int morton, uint32_t budget, uint32_t iter_built, uint32_t iter_locate, float ratio,
                           uint32_t spl, uint32_t cap, uint32_t* state, uint32_t* iter, uint32_t* pass, uint32_t* ev) {
  uint32_t per_all = 0, per_tmp = 0, cur = 0, sfp = 0;
  uint32_t remain = ((budget - 1u + spl) / spl) * spl;
  uint32_t n = 0;
  int executing = 1;
  while (executing && n < cap) {
    uint32_t e = 0;
    /* BegTrace */
    if (per_tmp == 0) {
      remain -= sfp;
      uint64_t want = (uint64_t)(1ull << (cur < 63u ? cur : 63u)) * spl;
      sfp = want < remain ? (uint32_t)want : remain;
      if ((remain - sfp < 2u * sfp) || ((float)per_all >= ratio * (float)budget)) sfp = remain;
      if (!morton && remain > sfp) e |= 1u;
    }
    uint32_t st;
    if (cur > iter_built) {
      st = remain > sfp ? 1u : 2u;
    } else if (!morton || cur > iter_locate) {
      st = 0u;
    } else {
      st = 3u;
    }
    state[n] = st;
    iter[n] = cur;
    pass[n] = sfp;
    /* EndTrace */
    per_all += spl;
    per_tmp += spl;
    if (per_tmp >= sfp) {
      e |= 2u;
      ++cur;
      if (morton && cur > iter_built) e |= 4u;
      per_tmp = 0;
    }
    if (morton && per_tmp == 0) e |= 8u;
    if (per_all > budget) {
      executing = 0;
      per_all = 0;
    }
    ev[n] = e;
    ++n;
  }
  return n;
}
