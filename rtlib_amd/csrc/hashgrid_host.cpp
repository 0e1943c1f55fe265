// Hash grid + Morton quadtree: host state, C ABI and pass controller.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/rtpg.h"
#include "context.h"
#include "hashgrid.h"

using namespace rtpg;

namespace {
template <class F>
int guarded(F&& f) {
  try {
    f();
    return RTPG_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RTPG_ERR_INVALID;
  } catch (...) {
    set_last_error("unknown error");
    return RTPG_ERR_INVALID;
  }
}
void require(bool c, const char* msg) {
  if (!c) throw Error(RTPG_ERR_INVALID, msg);
}
}  // namespace

// DoubleBufferedHashGrid3Buffer (RTLibExtOPX7Test.h:1472-1570).
struct rtpg_hashgrid_t {
  rtpg_context_t* ctx = nullptr;
  float origin[3], size[3];
  uint32_t bounds[3];
  uint32_t slots = 0;
  uint32_t key_mode = 0;
  DBuf<uint32_t> cur, prv;
  HashGridView view() {
    HashGridView v{};
    for (int k = 0; k < 3; ++k) {
      v.origin[k] = origin[k];
      v.size[k] = size[k];
      v.bounds[k] = bounds[k];
    }
    v.slots = slots;
    v.key_mode = key_mode;
    v.prv = prv.p;
    v.cur = cur.p;
    return v;
  }
};

// RTMortonQuadTreeWrapperT (OPX7UtilsMorton.h:590-764): two weight pyramids
// per hash slot, building and sampling, swapped at Update.
struct rtpg_mqtree_t {
  rtpg_context_t* ctx = nullptr;
  uint32_t slots = 0, level = 0, per_cell = 0;
  float fraction = 0.3f;
  DBuf<float> w[2];
  int building = 0;
  MortonView view() {
    MortonView m{};
    m.build = w[building].p;
    m.samp = w[1 - building].p;
    m.level = level;
    m.per_cell = per_cell;
    m.fraction = fraction;
    return m;
  }
};

namespace rtpg {
void guide_views(const rtpg_render_params* rp, rtpg_context_t* ctx, HashGridView& hg, MortonView& mq) {
  rtpg_hashgrid_t* g = rp->hashgrid;
  rtpg_mqtree_t* m = rp->mqtree;
  if (g->ctx != ctx || m->ctx != ctx) throw Error(RTPG_ERR_INVALID, "guide belongs to another context");
  if (m->slots != g->slots) throw Error(RTPG_ERR_INVALID, "mqtree slots must equal hash grid slots");
  hg = g->view();
  mq = m->view();
}
}  // namespace rtpg

extern "C" {

int rtpg_hashgrid_create(rtpg_context ctx, const float aabb_min[3], const float aabb_max[3], const uint32_t bounds[3],
                         uint32_t slots, uint32_t key_mode, rtpg_hashgrid* out) {
  return guarded([&] {
    require(ctx && aabb_min && aabb_max && bounds && out, "NULL argument");
    require(slots >= 32 && slots % 32 == 0, "slots must be a positive multiple of 32");
    require(key_mode <= 1, "key_mode must be 0 or 1");
    for (int k = 0; k < 3; ++k) require(bounds[k] >= 1, "bounds must be >= 1");
    if (key_mode == 1) {
      for (int k = 0; k < 3; ++k) require(bounds[k] <= 1024, "Morton keys need bounds <= 1024 per axis");
    }
    auto* g = new rtpg_hashgrid_t();
    g->ctx = ctx;
    for (int k = 0; k < 3; ++k) {
      g->origin[k] = aabb_min[k];
      g->size[k] = aabb_max[k] - aabb_min[k];
      g->bounds[k] = bounds[k];
    }
    g->slots = slots;
    g->key_mode = key_mode;
    try {
      g->cur.alloc(slots);
      g->prv.alloc(slots);
      g->cur.zero(slots, ctx->stream);
      g->prv.zero(slots, ctx->stream);
      RTPG_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      delete g;
      throw;
    }
    g->ctx->children++;
    *out = g;
  });
}

int rtpg_hashgrid_destroy(rtpg_hashgrid g) {
  return guarded([&] {
    if (!g) return;
    (void)hipStreamSynchronize(g->ctx->stream);
    g->ctx->children--;
    delete g;
  });
}

int rtpg_hashgrid_clear(rtpg_hashgrid g) {
  return guarded([&] {
    require(g != nullptr, "grid is NULL");
    g->cur.zero(g->slots, g->ctx->stream);
    g->prv.zero(g->slots, g->ctx->stream);
    RTPG_HIP(hipStreamSynchronize(g->ctx->stream));
  });
}

int rtpg_hashgrid_update(rtpg_hashgrid g) {
  return guarded([&] {
    require(g != nullptr, "grid is NULL");
    RTPG_HIP(hipMemcpyAsync(g->prv.p, g->cur.p, sizeof(uint32_t) * g->slots, hipMemcpyDeviceToDevice,
                            g->ctx->stream));
  });
}

int rtpg_hashgrid_find(rtpg_hashgrid g, const float* pos, uint32_t n, int insert, uint32_t* slot) {
  return guarded([&] {
    require(g && pos && slot, "NULL argument");
    if (n == 0) return;
    hipStream_t s = g->ctx->stream;
    DBuf<float> dp;
    DBuf<uint32_t> dout;
    dp.upload(pos, 3 * (size_t)n, s);
    dout.alloc(n);
    RTPG_HIP(launch_grid_find(g->view(), dp.p, n, insert != 0, dout.p, s));
    dout.download(slot, n, s);
  });
}

int rtpg_hashgrid_download(rtpg_hashgrid g, int which, uint32_t* checksums) {
  return guarded([&] {
    require(g && checksums, "NULL argument");
    (which ? g->prv : g->cur).download(checksums, g->slots, g->ctx->stream);
  });
}

int rtpg_mqtree_create(rtpg_context ctx, uint32_t slots, uint32_t level, float fraction, rtpg_mqtree* out) {
  return guarded([&] {
    require(ctx && out, "NULL argument");
    require(slots >= 1, "slots must be >= 1");
    require(level >= 1 && level <= 8, "level must be in [1, 8]");
    auto* m = new rtpg_mqtree_t();
    m->ctx = ctx;
    m->slots = slots;
    m->level = level;
    m->per_cell = ((1u << (2u * (level + 1u))) - 1u) / 3u;
    m->fraction = fraction;
    try {
      size_t n = (size_t)slots * m->per_cell;
      m->w[0].alloc(n);
      m->w[1].alloc(n);
      RTPG_HIP(launch_fill_zero(m->w[0].p, n, ctx->stream));
      RTPG_HIP(launch_fill_zero(m->w[1].p, n, ctx->stream));
      RTPG_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      delete m;
      throw;
    }
    m->ctx->children++;
    *out = m;
  });
}

int rtpg_mqtree_destroy(rtpg_mqtree m) {
  return guarded([&] {
    if (!m) return;
    (void)hipStreamSynchronize(m->ctx->stream);
    m->ctx->children--;
    delete m;
  });
}

int rtpg_mqtree_allocate(rtpg_mqtree m) {
  return guarded([&] {
    require(m != nullptr, "tree is NULL");
    size_t n = (size_t)m->slots * m->per_cell;
    RTPG_HIP(launch_fill_zero(m->w[0].p, n, m->ctx->stream));
    RTPG_HIP(launch_fill_zero(m->w[1].p, n, m->ctx->stream));
  });
}

int rtpg_mqtree_update(rtpg_mqtree m) {
  return guarded([&] {
    require(m != nullptr, "tree is NULL");
    m->building = 1 - m->building;
    RTPG_HIP(launch_fill_zero(m->w[m->building].p, (size_t)m->slots * m->per_cell, m->ctx->stream));
  });
}

int rtpg_mqtree_build(rtpg_mqtree m) {
  return guarded([&] {
    require(m != nullptr, "tree is NULL");
    RTPG_HIP(launch_morton_build(m->w[1 - m->building].p, m->level, m->per_cell, m->slots, m->ctx->stream));
  });
}

int rtpg_mqtree_record(rtpg_mqtree m, const uint32_t* slot, const float* canon, const float* value, uint32_t n) {
  return guarded([&] {
    require(m && slot && canon && value, "NULL argument");
    if (n == 0) return;
    for (uint32_t i = 0; i < n; ++i) require(slot[i] == 0xFFFFFFFFu || slot[i] < m->slots, "slot out of range");
    hipStream_t s = m->ctx->stream;
    DBuf<uint32_t> dc;
    DBuf<float> dq, dv;
    dc.upload(slot, n, s);
    dq.upload(canon, 2 * (size_t)n, s);
    dv.upload(value, n, s);
    RTPG_HIP(launch_morton_record(m->view(), dc.p, dq.p, dv.p, n, s));
    RTPG_HIP(hipStreamSynchronize(s));
  });
}

int rtpg_mqtree_sample(rtpg_mqtree m, const uint32_t* slot, uint32_t n, uint64_t seed, float* out) {
  return guarded([&] {
    require(m && slot && out, "NULL argument");
    if (n == 0) return;
    for (uint32_t i = 0; i < n; ++i) require(slot[i] < m->slots, "slot out of range");
    hipStream_t s = m->ctx->stream;
    DBuf<uint32_t> dc;
    DBuf<float> dout;
    dc.upload(slot, n, s);
    dout.alloc(3 * (size_t)n);
    RTPG_HIP(launch_morton_sample(m->view(), dc.p, n, seed, dout.p, s));
    dout.download(out, 3 * (size_t)n, s);
  });
}

int rtpg_mqtree_pdf(rtpg_mqtree m, const uint32_t* slot, const float* canon, uint32_t n, float* pdf) {
  return guarded([&] {
    require(m && slot && canon && pdf, "NULL argument");
    if (n == 0) return;
    for (uint32_t i = 0; i < n; ++i) require(slot[i] < m->slots, "slot out of range");
    hipStream_t s = m->ctx->stream;
    DBuf<uint32_t> dc;
    DBuf<float> dq, dout;
    dc.upload(slot, n, s);
    dq.upload(canon, 2 * (size_t)n, s);
    dout.alloc(n);
    RTPG_HIP(launch_morton_pdf(m->view(), dc.p, dq.p, n, dout.p, s));
    dout.download(pdf, n, s);
  });
}

int rtpg_mqtree_download(rtpg_mqtree m, int which, float* weights) {
  return guarded([&] {
    require(m && weights, "NULL argument");
    int idx = which ? 1 - m->building : m->building;
    m->w[idx].download(weights, (size_t)m->slots * m->per_cell, m->ctx->stream);
  });
}

int rtpg_mqtree_upload(rtpg_mqtree m, int which, const float* weights) {
  return guarded([&] {
    require(m && weights, "NULL argument");
    int idx = which ? 1 - m->building : m->building;
    m->w[idx].upload(weights, (size_t)m->slots * m->per_cell, m->ctx->stream);
    RTPG_HIP(hipStreamSynchronize(m->ctx->stream));
  });
}

int rtpg_mqtree_info(rtpg_mqtree m, uint32_t info[3]) {
  return guarded([&] {
    require(m && info, "NULL argument");
    info[0] = m->slots;
    info[1] = m->level;
    info[2] = m->per_cell;
  });
}

}  // extern "C"

// RTMortonQuadTreeControllerT (OPX7UtilsMorton.h:766-993) with the hash grid
// snapshot of the application (RTLibExtOPX7TestApplication.cpp:2208-2222).
struct rtpg_mqtree_ctrl_t {
  rtpg_mqtree_t* tree = nullptr;
  rtpg_hashgrid_t* grid = nullptr;
  uint32_t budget = 0, spl = 1, iter_built = 0, iter_locate = 0;
  float ratio = 0.5f;
  bool start = false, executing = false;
  uint32_t per_all = 0, per_tmp = 0, remain = 0, pass = 0, iter = 0;
  uint32_t state = 0;
};

extern "C" {

int rtpg_mqtree_ctrl_create(rtpg_mqtree tree, rtpg_hashgrid grid, uint32_t sample_for_budget,
                            uint32_t iteration_for_built, uint32_t iteration_for_locate, float ratio_for_budget,
                            uint32_t sample_per_launch, rtpg_mqtree_ctrl* out) {
  return guarded([&] {
    require(tree && out, "NULL argument");
    require(sample_per_launch >= 1 && sample_for_budget >= 1, "budget and samples per launch must be >= 1");
    auto* c = new rtpg_mqtree_ctrl_t();
    c->tree = tree;
    c->grid = grid;
    c->budget = sample_for_budget;
    c->iter_built = iteration_for_built;
    c->iter_locate = iteration_for_locate;
    c->ratio = ratio_for_budget;
    c->spl = sample_per_launch;
    *out = c;
  });
}

int rtpg_mqtree_ctrl_destroy(rtpg_mqtree_ctrl c) {
  delete c;
  return RTPG_OK;
}

int rtpg_mqtree_ctrl_start(rtpg_mqtree_ctrl c) {
  return guarded([&] {
    require(c != nullptr, "ctrl is NULL");
    c->start = true;
  });
}

// state: 0 Record, 1 RecordAndSample, 2 Sample, 3 Locate.
int rtpg_mqtree_ctrl_beg_trace(rtpg_mqtree_ctrl c) {
  return guarded([&] {
    require(c != nullptr, "ctrl is NULL");
    if (c->start) {
      c->per_all = 0;
      c->per_tmp = 0;
      c->iter = 0;
      c->start = false;
      c->executing = true;
      c->remain = ((c->budget - 1 + c->spl) / c->spl) * c->spl;
      c->pass = 0;
      c->state = 0;
      size_t n = (size_t)c->tree->slots * c->tree->per_cell;
      RTPG_HIP(launch_fill_zero(c->tree->w[0].p, n, c->tree->ctx->stream));
      RTPG_HIP(launch_fill_zero(c->tree->w[1].p, n, c->tree->ctx->stream));
      if (c->grid) {
        c->grid->cur.zero(c->grid->slots, c->grid->ctx->stream);
        c->grid->prv.zero(c->grid->slots, c->grid->ctx->stream);
      }
    }
    if (!c->executing) return;
    if (c->per_tmp == 0) {
      c->remain -= c->pass;
      uint64_t want = (uint64_t)(1ull << std::min(c->iter, 31u)) * c->spl;
      c->pass = (uint32_t)std::min<uint64_t>(c->remain, want);
      if ((c->remain - c->pass < 2 * c->pass) || ((float)c->per_all >= c->ratio * (float)c->budget)) {
        c->pass = c->remain;
      }
    }
    if (c->iter > c->iter_built) {
      c->state = c->remain > c->pass ? 1u : 2u;
    } else if (c->iter > c->iter_locate) {
      c->state = 0u;
    } else {
      c->state = 3u;
    }
  });
}

int rtpg_mqtree_ctrl_end_trace(rtpg_mqtree_ctrl c, uint32_t samples) {
  return guarded([&] {
    require(c != nullptr, "ctrl is NULL");
    if (!c->executing) return;
    uint32_t k = samples ? samples : c->spl;
    c->per_all += k;
    c->per_tmp += k;
    if (c->per_tmp >= c->pass) {
      c->iter++;
      if (c->iter > c->iter_built) {
        rtpg_mqtree_t* m = c->tree;
        m->building = 1 - m->building;
        RTPG_HIP(launch_fill_zero(m->w[m->building].p, (size_t)m->slots * m->per_cell, m->ctx->stream));
        RTPG_HIP(launch_morton_build(m->w[1 - m->building].p, m->level, m->per_cell, m->slots, m->ctx->stream));
      }
      c->per_tmp = 0;
      if (c->grid) {
        RTPG_HIP(hipMemcpyAsync(c->grid->prv.p, c->grid->cur.p, sizeof(uint32_t) * c->grid->slots,
                                hipMemcpyDeviceToDevice, c->grid->ctx->stream));
      }
    }
    if (c->per_all > c->budget) {
      c->executing = false;
      c->per_all = 0;
    }
  });
}

int rtpg_mqtree_ctrl_query(rtpg_mqtree_ctrl c, uint32_t q[6]) {
  return guarded([&] {
    require(c && q, "NULL argument");
    q[0] = c->state;
    q[1] = c->iter;
    q[2] = c->per_tmp;
    q[3] = c->pass;
    q[4] = c->executing ? 1u : 0u;
    q[5] = c->per_all;
  });
}

}  // extern "C"
