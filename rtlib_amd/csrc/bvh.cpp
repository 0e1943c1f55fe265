// Binned-SAH binary build collapsed into a 4-wide BVH (host C++, threaded).
#include "bvh.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <stdexcept>

namespace rtpg {
namespace {

struct Box {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(const float* p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.lo[k]);
      hi[k] = std::max(hi[k], b.hi[k]);
    }
  }
  float area() const {
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = std::max(hi[k] - lo[k], 0.0f);
    return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }
  bool valid() const { return lo[0] <= hi[0]; }
};

// binary16 bits of v >= 0 rounded down (up = false) or up, then one more
// binary16 step in the same direction.  frexp splits v exactly, and the
// mantissa scaled to 10 fraction bits is exact in double.
uint16_t half_outward(double v, bool up) {
  uint32_t b;
  if (v <= 0.0) {
    b = 0u;
  } else if (v >= 65504.0) {
    b = up && v > 65504.0 ? kHalfInf : 0x7BFFu;
  } else {
    int e;
    double m = std::frexp(v, &e);  // v = m 2^e, m in [0.5, 1)
    int E = e - 1;                 // v = (2m) 2^E, 2m in [1, 2)
    if (E < -14) {
      double q = std::ldexp(v, 24);  // subnormal: multiples of 2^-24
      b = (uint32_t)(up ? std::ceil(q) : std::floor(q));
    } else {
      double f = (2.0 * m - 1.0) * 1024.0;
      uint32_t fi = (uint32_t)(up ? std::ceil(f) : std::floor(f));
      b = ((uint32_t)(E + 15) << 10) + fi;  // fi = 1024 carries into the exponent
    }
  }
  if (up) return (uint16_t)(b >= kHalfInf ? kHalfInf : b + 1u);
  return (uint16_t)(b == 0u ? 0x8001u : b - 1u);  // 0x8001: -2^-24
}

struct PrimRef {
  Box box;
  float c[3];
  uint32_t id;
};

struct BinNode {
  Box box;
  int32_t left = -1, right = -1;  // children in the binary tree
  uint32_t start = 0, count = 0;  // leaf range in the prim-ref array
  bool leaf() const { return left < 0; }
};

constexpr int kBins = 32;
constexpr float kTravCost = 1.0f;
constexpr float kIsectCost = 1.0f;

struct Builder {
  std::vector<PrimRef>& refs;
  explicit Builder(std::vector<PrimRef>& r) : refs(r) {}

  // Splits [start, start+count) in place; returns the size of the left part,
  // or 0 when a leaf is the better choice.
  uint32_t split(uint32_t start, uint32_t count, const Box& box) {
    if (count <= 2) return 0;
    Box cb;
    for (uint32_t i = start; i < start + count; ++i) cb.grow(refs[i].c);
    float best_cost = FLT_MAX;
    int best_axis = -1, best_bin = -1;
    for (int axis = 0; axis < 3; ++axis) {
      float ext = cb.hi[axis] - cb.lo[axis];
      if (!(ext > 0.0f)) continue;
      Box bb[kBins];
      uint32_t bc[kBins] = {0};
      float scale = kBins / ext;
      for (uint32_t i = start; i < start + count; ++i) {
        int b = std::min(kBins - 1, (int)((refs[i].c[axis] - cb.lo[axis]) * scale));
        bb[b].grow(refs[i].box);
        bc[b]++;
      }
      float ra[kBins];
      uint32_t rc[kBins];
      Box acc;
      uint32_t n = 0;
      for (int b = kBins - 1; b > 0; --b) {
        acc.grow(bb[b]);
        n += bc[b];
        ra[b] = acc.valid() ? acc.area() : 0.0f;
        rc[b] = n;
      }
      Box lacc;
      uint32_t ln = 0;
      for (int b = 0; b < kBins - 1; ++b) {
        lacc.grow(bb[b]);
        ln += bc[b];
        if (ln == 0 || rc[b + 1] == 0) continue;
        float cost = (lacc.valid() ? lacc.area() : 0.0f) * ln + ra[b + 1] * rc[b + 1];
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_bin = b;
        }
      }
    }
    float parent_area = std::max(box.area(), 1e-30f);
    float leaf_cost = kIsectCost * count;
    float split_cost = kTravCost + kIsectCost * best_cost / parent_area;
    if (best_axis < 0) {
      // All centroids coincide: split by index if the leaf would be too big.
      if (count <= kMaxLeafTris) return 0;
      return count / 2;
    }
    if (count <= kMaxLeafTris && leaf_cost <= split_cost) return 0;
    float ext = cb.hi[best_axis] - cb.lo[best_axis];
    float scale = kBins / ext;
    auto mid = std::partition(refs.begin() + start, refs.begin() + start + count,
                              [&](const PrimRef& r) {
                                int b = std::min(kBins - 1, (int)((r.c[best_axis] - cb.lo[best_axis]) * scale));
                                return b <= best_bin;
                              });
    uint32_t nl = (uint32_t)(mid - (refs.begin() + start));
    if (nl == 0 || nl == count) return count / 2;
    return nl;
  }

  void build(std::vector<BinNode>& out, int32_t node) {
    // Iterative to keep deep trees off the host stack.
    std::vector<int32_t> work{node};
    while (!work.empty()) {
      int32_t ni = work.back();
      work.pop_back();
      uint32_t start = out[ni].start, count = out[ni].count;
      Box box = out[ni].box;
      uint32_t nl = split(start, count, box);
      if (nl == 0) continue;
      BinNode l, r;
      l.start = start;
      l.count = nl;
      r.start = start + nl;
      r.count = count - nl;
      for (uint32_t i = l.start; i < l.start + l.count; ++i) l.box.grow(refs[i].box);
      for (uint32_t i = r.start; i < r.start + r.count; ++i) r.box.grow(refs[i].box);
      int32_t li = (int32_t)out.size();
      out.push_back(l);
      out.push_back(r);
      out[ni].left = li;
      out[ni].right = li + 1;
      work.push_back(li + 1);
      work.push_back(li);
    }
  }
};

struct Collapser {
  const std::vector<BinNode>& bin;
  const std::vector<PrimRef>& refs;
  BvhBuildResult& res;
  Collapser(const std::vector<BinNode>& b, const std::vector<PrimRef>& r, BvhBuildResult& o)
      : bin(b), refs(r), res(o) {}

  // Emits the BVH4 node for binary node b (DFS order) and returns its index.
  int32_t emit(int32_t b, uint32_t depth, uint32_t stack_here) {
    int32_t me = (int32_t)res.nodes.size();
    res.nodes.emplace_back();
    res.max_depth = std::max(res.max_depth, depth);
    int32_t kids[4];
    int nk = 0;
    kids[nk++] = bin[b].left;
    kids[nk++] = bin[b].right;
    while (nk < 4) {
      int best = -1;
      float best_area = -1.0f;
      for (int i = 0; i < nk; ++i) {
        if (!bin[kids[i]].leaf()) {
          float a = bin[kids[i]].box.area();
          if (a > best_area) {
            best_area = a;
            best = i;
          }
        }
      }
      if (best < 0) break;
      int32_t k = kids[best];
      kids[best] = bin[k].left;
      kids[nk++] = bin[k].right;
    }
    // The traversal keeps every hit child but the nearest on its stack, leaves
    // included, so below this node at most nk - 1 more entries are stacked.
    uint32_t child_stack = stack_here + (uint32_t)(nk - 1);
    res.max_stack = std::max(res.max_stack, child_stack);
    int32_t child_ref[4];
    uint32_t child_cnt[4];
    for (int i = 0; i < nk; ++i) {
      const BinNode& c = bin[kids[i]];
      if (c.leaf()) {
        child_ref[i] = ~(int32_t)c.start;
        child_cnt[i] = c.count;
        res.max_leaf = std::max(res.max_leaf, c.count);
      } else {
        child_ref[i] = emit(kids[i], depth + 1, child_stack);
        child_cnt[i] = 0;
      }
    }
    BvhNode4& n = res.nodes[me];
    bool used[4];
    float org[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    for (int i = 0; i < 4; ++i) {
      // an empty leaf (the tiny-scene path) is an empty slot
      used[i] = i < nk && !(child_ref[i] < 0 && child_cnt[i] == 0);
      if (used[i])
        for (int k = 0; k < 3; ++k) org[k] = std::min(org[k], bin[kids[i]].box.lo[k]);
    }
    for (int k = 0; k < 3; ++k) n.org[k] = org[k] == FLT_MAX ? 0.0f : org[k];
    n.pad = 0;
    uint16_t* planes[3] = {n.x, n.y, n.z};
    for (int i = 0; i < 4; ++i) {
      if (used[i]) {
        const Box& bx = bin[kids[i]].box;
        for (int k = 0; k < 3; ++k) {
          planes[k][i] = half_outward((double)bx.lo[k] - (double)n.org[k], false);
          planes[k][4 + i] = half_outward((double)bx.hi[k] - (double)n.org[k], true);
        }
        n.entry[i] = child_ref[i] >= 0 ? (uint32_t)child_ref[i]
                                       : (kLeafEntry | ((child_cnt[i] - 1u) << 27) | (uint32_t)(~child_ref[i]));
      } else {
        for (int k = 0; k < 3; ++k) planes[k][i] = planes[k][4 + i] = kHalfInf;
        n.entry[i] = 0;
      }
    }
    return me;
  }
};

}  // namespace

BvhBuildResult build_bvh4(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                          int threads) {
  if (nt == 0) throw std::runtime_error("scene has no triangles");
  std::vector<PrimRef> refs(nt);
  for (uint32_t t = 0; t < nt; ++t) {
    PrimRef& r = refs[t];
    for (int k = 0; k < 3; ++k) {
      uint32_t vi = idx[3 * t + k];
      if (vi >= nv) throw std::runtime_error("triangle index out of range");
      r.box.grow(verts + 3 * vi);
    }
    for (int k = 0; k < 3; ++k) r.c[k] = 0.5f * (r.box.lo[k] + r.box.hi[k]);
    r.id = t;
  }
  // Root, then the top of the tree serially until there are enough subtrees
  // for the worker threads.
  std::vector<BinNode> top;
  top.emplace_back();
  top[0].start = 0;
  top[0].count = nt;
  for (auto& r : refs) top[0].box.grow(r.box);
  Builder bld(refs);
  std::vector<int32_t> frontier{0};
  int want = std::max(1, threads) * 4;
  while ((int)frontier.size() < want) {
    std::vector<int32_t> next;
    bool any = false;
    for (int32_t ni : frontier) {
      uint32_t nl = bld.split(top[ni].start, top[ni].count, top[ni].box);
      if (nl == 0) continue;
      any = true;
      BinNode l, r;
      l.start = top[ni].start;
      l.count = nl;
      r.start = top[ni].start + nl;
      r.count = top[ni].count - nl;
      for (uint32_t i = l.start; i < l.start + l.count; ++i) l.box.grow(refs[i].box);
      for (uint32_t i = r.start; i < r.start + r.count; ++i) r.box.grow(refs[i].box);
      int32_t li = (int32_t)top.size();
      top.push_back(l);
      top.push_back(r);
      top[ni].left = li;
      top[ni].right = li + 1;
      next.push_back(li);
      next.push_back(li + 1);
    }
    if (!any) break;
    frontier.swap(next);
  }
  // Subtrees in parallel, each into its own node vector (root at index 0).
  std::vector<std::vector<BinNode>> sub(frontier.size());
  {
    std::vector<std::thread> pool;
    int nth = std::max(1, threads);
    std::vector<size_t> jobs(frontier.size());
    for (size_t i = 0; i < jobs.size(); ++i) jobs[i] = i;
    for (int t = 0; t < nth; ++t) {
      pool.emplace_back([&, t]() {
        for (size_t j = (size_t)t; j < jobs.size(); j += (size_t)nth) {
          std::vector<BinNode>& v = sub[j];
          v.push_back(top[frontier[j]]);
          v[0].left = v[0].right = -1;
          Builder b(refs);
          b.build(v, 0);
        }
      });
    }
    for (auto& th : pool) th.join();
  }
  // Merge into one binary tree.
  std::vector<BinNode> bin = top;
  for (size_t j = 0; j < frontier.size(); ++j) {
    int32_t base = (int32_t)bin.size();
    const std::vector<BinNode>& v = sub[j];
    bin[frontier[j]] = v[0];
    if (!v[0].leaf()) {
      bin[frontier[j]].left = v[0].left - 1 + base;
      bin[frontier[j]].right = v[0].right - 1 + base;
    }
    for (size_t k = 1; k < v.size(); ++k) {
      BinNode n = v[k];
      if (!n.leaf()) {
        n.left = n.left - 1 + base;
        n.right = n.right - 1 + base;
      }
      bin.push_back(n);
    }
  }
  BvhBuildResult res;
  for (int k = 0; k < 3; ++k) {
    res.world_lo[k] = bin[0].box.lo[k];
    res.world_hi[k] = bin[0].box.hi[k];
  }
  if (bin[0].leaf()) {
    // Tiny scene: a root with one leaf child.
    BinNode l = bin[0];
    bin.push_back(l);
    BinNode e;
    e.start = l.start + l.count;
    e.count = 0;
    bin.push_back(e);
    bin[0].left = 1;
    bin[0].right = 2;
    bin[2].box = Box();
  }
  Collapser col(bin, refs, res);
  col.emit(0, 1, 0);
  res.tris.resize(nt);
  res.order.resize(nt);
  for (uint32_t s = 0; s < nt; ++s) {
    uint32_t t = refs[s].id;
    res.order[s] = t;
    const float* a = verts + 3 * idx[3 * t + 0];
    const float* b = verts + 3 * idx[3 * t + 1];
    const float* c = verts + 3 * idx[3 * t + 2];
    TriRec& tr = res.tris[s];
    uint32_t idbits = t;
    for (int k = 0; k < 3; ++k) {
      tr.v0[k] = a[k];
      tr.e1[k] = b[k] - a[k];
      tr.e2[k] = c[k] - a[k];
    }
    std::memcpy(&tr.v0[3], &idbits, 4);
    tr.e1[3] = 0.0f;
    tr.e2[3] = 0.0f;
  }
  return res;
}

}  // namespace rtpg
