// C-ABI of the learned directional guide: allocation, initialisation,
// queries, gradient and Adam training steps.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rtpg.h"
#include "context.h"
#include "mlp.h"

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

rtpg_mlp_config default_config() {
  rtpg_mlp_config c{};
  c.log2_table = 19;
  c.base_res = 16.0f;
  c.max_res = 256.0f;
  c.lr = 1e-2f;
  c.beta1 = 0.9f;
  c.beta2 = 0.99f;
  c.eps = 1e-15f;
  c.batch = 1u << 18;
  c.capacity = 1u << 20;
  c.record_paths = 1u << 19;
  c.seed = 0;
  c.guide_depth = 0;
  c.cache_res = 0;
  return c;
}

// Packs the host sample arrays into the device layout (four float4 per sample).
void upload_samples(rtpg_mlp_t* m, const float* pos, const float* wo, const float* nrm, const float* rough,
                    const float* omega, const float* weight, const float* pdf_bsdf, uint32_t n) {
  std::vector<float4> h(4 * (size_t)n);
  for (size_t i = 0; i < n; ++i) {
    h[4 * i + 0] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], rough[i]);
    h[4 * i + 1] = make_float4(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], pdf_bsdf[i]);
    h[4 * i + 2] = make_float4(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2], weight[i]);
    h[4 * i + 3] = make_float4(omega[3 * i], omega[3 * i + 1], omega[3 * i + 2], 0.0f);
  }
  m->staging.upload(h, m->ctx->stream);
}

uint32_t train_grid(uint32_t n) { return std::max(1u, std::min(kMlpTrainGrid, (n + 63u) / 64u)); }

// Forward + backward over n device samples; leaves the weight gradient sum in
// w_grad, the table gradient sum in t_grad and the loss sum in loss[0].
// The loss sum goes to loss[slot] (loss holds one slot per step of a call).
void grad_pass(rtpg_mlp_t* m, const float4* smp, uint32_t n, uint32_t slot = 0) {
  hipStream_t s = m->ctx->stream;
  uint32_t grid = train_grid(n);
  m->loss.reserve(slot + 1u);
  RTPG_HIP(launch_mlp_train(mlp_view_of(m, true), smp, n, m->partial.p, m->loss_partial.p, grid, s));
  RTPG_HIP(launch_mlp_reduce(m->partial.p, m->loss_partial.p, grid, m->w_grad.p, m->loss.p + slot, s));
}

void adam_step(rtpg_mlp_t* m, uint32_t n) {
  hipStream_t s = m->ctx->stream;
  m->steps += 1;
  m->cache_stale = true;
  const rtpg_mlp_config& c = m->cfg;
  double b1t = std::pow((double)c.beta1, (double)m->steps), b2t = std::pow((double)c.beta2, (double)m->steps);
  float lr_t = (float)(c.lr * std::sqrt(1.0 - b2t) / (1.0 - b1t));
  float gs = 1.0f / (float)n;
  RTPG_HIP(launch_mlp_adam(m->w.p, m->w_m1.p, m->w_m2.p, m->w_grad.p, gs, kMlpWeights, lr_t, c.beta1, c.beta2, c.eps,
                           m->w_bf16.p, nullptr, nullptr, 0, s));
  RTPG_HIP(launch_mlp_adam(m->t.p, m->t_m1.p, m->t_m2.p, m->t_grad.p, gs, m->table_floats(), lr_t, c.beta1, c.beta2,
                           c.eps, nullptr, m->t_bf16.p, m->t_grad_pk.p, (size_t)kMlpPkFirst * m->entries, s));
}

float download_loss(rtpg_mlp_t* m, uint32_t n) {
  float l = 0.0f;
  m->loss.download(&l, 1, m->ctx->stream);
  return l / (float)n;
}

void reset_state(rtpg_mlp_t* m, uint64_t seed) {
  hipStream_t s = m->ctx->stream;
  m->cache_stale = true;
  size_t tf = m->table_floats();
  RTPG_HIP(launch_mlp_init(m->w.p, m->w_bf16.p, m->t.p, m->t_bf16.p, tf, seed, s));
  m->w_m1.zero(kMlpWeights, s);
  m->w_m2.zero(kMlpWeights, s);
  m->w_grad.zero(kMlpWeights, s);
  m->t_m1.zero(tf, s);
  m->t_m2.zero(tf, s);
  m->t_grad.zero(tf, s);
  m->t_grad_pk.zero(m->pk_entries(), s);
  m->count.zero(1, s);
  m->steps = 0;
  RTPG_HIP(hipStreamSynchronize(s));
}

void check_samples(const float* pos, const float* wo, const float* nrm, const float* rough, const float* omega,
                   const float* weight, const float* pdf_bsdf) {
  require(pos && wo && nrm && rough && omega && weight && pdf_bsdf, "NULL sample array");
}

}  // namespace

MlpView mlp_view_of(const rtpg_mlp_t* m, bool training) {
  MlpView v{};
  v.table = m->t_bf16.p;
  v.grad = training ? m->t_grad.p : nullptr;
  v.grad_pk = training ? m->t_grad_pk.p : nullptr;
  v.w = m->w_bf16.p;
  for (int k = 0; k < 3; ++k) v.lo[k] = m->lo[k];
  v.inv_size = 1.0f / m->size;
  v.tmask = (uint32_t)(m->entries - 1);
  v.dense = m->dense;
  for (int l = 0; l < kMlpLevels; ++l) v.res[l] = m->res[l];
  v.ready = 1u;
  v.cache_res = m->cfg.cache_res;
  return v;
}

void mlp_cache_refresh(rtpg_mlp_t* m) {
  if (m->cfg.cache_res == 0 || !m->cache_stale) return;
  uint32_t n = m->cache_entries();
  m->cache_mix.reserve((size_t)kMlpMixF4 * n);
  m->cache_c.reserve(n);
  RTPG_HIP(launch_mlp_cache_fill(mlp_view_of(m, false), n, m->cache_mix.p, m->cache_c.p, m->ctx->stream));
  m->cache_stale = false;
}

extern "C" {

int rtpg_mlp_create(rtpg_context ctx, const float aabb_min[3], const float aabb_max[3], const rtpg_mlp_config* cfg,
                    rtpg_mlp* out) {
  return guarded([&] {
    require(ctx && aabb_min && aabb_max && out, "NULL argument");
    rtpg_mlp_config c = cfg ? *cfg : default_config();
    require(c.log2_table >= 8 && c.log2_table <= 24, "log2_table must be in [8, 24]");
    require(c.base_res >= 2.0f && c.max_res >= c.base_res && c.max_res <= 1e6f, "need 2 <= base_res <= max_res");
    require(c.lr > 0.0f && c.beta1 >= 0.0f && c.beta1 < 1.0f && c.beta2 >= 0.0f && c.beta2 < 1.0f && c.eps >= 0.0f,
            "bad Adam parameters");
    require(c.batch >= 1 && c.capacity >= 1 && c.record_paths >= 1, "batch, capacity and record_paths must be >= 1");
    require(c.cache_res <= 256, "cache_res must be in [0, 256]");
    RTPG_HIP(hipSetDevice(ctx->device));
    auto m = new rtpg_mlp_t();
    try {
      m->ctx = ctx;
      m->cfg = c;
      float ext = 0.0f;
      for (int k = 0; k < 3; ++k) {
        require(aabb_max[k] >= aabb_min[k], "empty box");
        ext = std::max(ext, aabb_max[k] - aabb_min[k]);
      }
      ext = std::max(ext, 1e-6f);
      for (int k = 0; k < 3; ++k) m->lo[k] = 0.5f * (aabb_min[k] + aabb_max[k]) - 0.5f * ext;
      m->size = ext;
      m->entries = (size_t)1 << c.log2_table;
      double b = std::exp((std::log((double)c.max_res) - std::log((double)c.base_res)) / (kMlpLevels - 1));
      m->dense = 0;
      for (int l = 0; l < kMlpLevels; ++l) {
        double r = std::floor((double)c.base_res * std::pow(b, (double)l) + 1e-9);
        m->res[l] = (float)r;
        double cells = (r + 1.0) * (r + 1.0) * (r + 1.0);
        if (cells <= (double)m->entries) m->dense |= 1u << l;
      }
      size_t tf = m->table_floats();
      m->w.alloc(kMlpWeights);
      m->w_m1.alloc(kMlpWeights);
      m->w_m2.alloc(kMlpWeights);
      m->w_grad.alloc(kMlpWeights);
      m->w_bf16.alloc(kMlpWeights);
      m->t.alloc(tf);
      m->t_m1.alloc(tf);
      m->t_m2.alloc(tf);
      m->t_grad.alloc(tf);
      m->t_grad_pk.alloc(m->pk_entries());
      m->t_bf16.alloc(tf / 2);
      m->partial.alloc((size_t)kMlpTrainGrid * kMlpWeights);
      m->loss_partial.alloc(kMlpTrainGrid);
      m->loss.alloc(1);
      m->samples.alloc(4 * (size_t)c.capacity);
      m->count.alloc(1);
      reset_state(m, c.seed);
    } catch (...) {
      delete m;
      throw;
    }
    m->ctx->children++;
    *out = m;
  });
}

int rtpg_mlp_destroy(rtpg_mlp m) {
  return guarded([&] {
    if (m) {
      (void)hipSetDevice(m->ctx->device);
      (void)hipStreamSynchronize(m->ctx->stream);
      m->ctx->children--;
    }
    delete m;
  });
}

int rtpg_mlp_reset(rtpg_mlp m, uint64_t seed) {
  return guarded([&] {
    require(m != nullptr, "NULL mlp");
    RTPG_HIP(hipSetDevice(m->ctx->device));
    reset_state(m, seed);
  });
}

int rtpg_mlp_forward(rtpg_mlp m, const float* pos, const float* wo, const float* nrm, const float* rough, uint32_t n,
                     float* raw, float* mixture) {
  return guarded([&] {
    require(m && pos && wo && nrm && rough, "NULL argument");
    if (n == 0) return;
    RTPG_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    std::vector<float4> q(3 * (size_t)n);
    for (size_t i = 0; i < n; ++i) {
      q[3 * i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], rough[i]);
      q[3 * i + 1] = make_float4(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 0.0f);
      q[3 * i + 2] = make_float4(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2], 0.0f);
    }
    m->staging.upload(q, s);
    size_t per = (size_t)kMlpOutUsed * n;
    m->out.reserve(2 * per);
    RTPG_HIP(launch_mlp_forward(mlp_view_of(m, false), m->staging.p, n, raw ? m->out.p : nullptr,
                                mixture ? m->out.p + per : nullptr, s));
    if (raw) m->out.download(raw, per, s);
    if (mixture) RTPG_HIP(hipMemcpyAsync(mixture, m->out.p + per, per * sizeof(float), hipMemcpyDeviceToHost, s));
    RTPG_HIP(hipStreamSynchronize(s));
  });
}

int rtpg_mlp_sample(rtpg_mlp m, const float* mixture, uint32_t n, uint64_t seed, float* out) {
  return guarded([&] {
    require(m && mixture && out, "NULL argument");
    if (n == 0) return;
    RTPG_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    size_t per = (size_t)kMlpOutUsed * n;
    m->out.upload(mixture, per, s);
    m->staging.reserve(n);
    RTPG_HIP(launch_mlp_sample(m->out.p, n, seed, m->staging.p, s));
    std::vector<float4> h = m->staging.to_vector(n, s);
    std::memcpy(out, h.data(), sizeof(float4) * n);
  });
}

int rtpg_mlp_gradient(rtpg_mlp m, const float* pos, const float* wo, const float* nrm, const float* rough,
                      const float* omega, const float* weight, const float* pdf_bsdf, uint32_t n, float* loss,
                      float* grad_w, float* grad_table) {
  return guarded([&] {
    require(m != nullptr, "NULL mlp");
    check_samples(pos, wo, nrm, rough, omega, weight, pdf_bsdf);
    require(n >= 1, "n must be >= 1");
    RTPG_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    upload_samples(m, pos, wo, nrm, rough, omega, weight, pdf_bsdf, n);
    m->t_grad.zero(m->table_floats(), s);
    m->t_grad_pk.zero(m->pk_entries(), s);
    grad_pass(m, m->staging.p, n);
    float inv = 1.0f / (float)n;
    if (loss) *loss = download_loss(m, n);
    if (grad_w) {
      m->w_grad.download(grad_w, kMlpWeights, s);
      for (int k = 0; k < kMlpWeights; ++k) grad_w[k] *= inv;
    }
    if (grad_table) {
      size_t tf = m->table_floats();
      m->t_grad.download(grad_table, tf, s);
      // fine levels: the packed bf16 pairs
      std::vector<uint32_t> pk = m->t_grad_pk.to_vector(m->pk_entries(), s);
      float* fine = grad_table + (size_t)kMlpPkFirst * m->entries * kMlpFeat;
      for (size_t e = 0; e < pk.size(); ++e) {
        fine[2 * e] += bf2f((uint16_t)(pk[e] & 0xFFFFu));
        fine[2 * e + 1] += bf2f((uint16_t)(pk[e] >> 16));
      }
      for (size_t k = 0; k < tf; ++k) grad_table[k] *= inv;
    }
    m->t_grad.zero(m->table_floats(), s);
    m->t_grad_pk.zero(m->pk_entries(), s);
    RTPG_HIP(hipStreamSynchronize(s));
  });
}

int rtpg_mlp_train(rtpg_mlp m, const float* pos, const float* wo, const float* nrm, const float* rough,
                   const float* omega, const float* weight, const float* pdf_bsdf, uint32_t n, uint32_t steps,
                   float* loss) {
  return guarded([&] {
    require(m != nullptr, "NULL mlp");
    check_samples(pos, wo, nrm, rough, omega, weight, pdf_bsdf);
    require(n >= 1, "n must be >= 1");
    RTPG_HIP(hipSetDevice(m->ctx->device));
    upload_samples(m, pos, wo, nrm, rough, omega, weight, pdf_bsdf, n);
    for (uint32_t k = 0; k < steps; ++k) {
      grad_pass(m, m->staging.p, n);
      if (loss) loss[k] = download_loss(m, n);
      adam_step(m, n);
    }
    RTPG_HIP(hipStreamSynchronize(m->ctx->stream));
  });
}

int rtpg_mlp_train_recorded(rtpg_mlp m, uint32_t steps, float* loss, uint32_t* used) {
  return guarded([&] {
    require(m != nullptr, "NULL mlp");
    RTPG_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    uint32_t cnt = 0;
    m->count.download(&cnt, 1, s);
    uint32_t n = std::min(cnt, m->cfg.capacity);
    if (used) *used = n;
    if (n == 0) {
      if (loss)
        for (uint32_t k = 0; k < steps; ++k) loss[k] = 0.0f;
      return;
    }
    uint32_t b = std::min(n, m->cfg.batch);
    uint32_t windows = n / b;  // whole batches in recording order
    // each window the steps use, copied in spatial order (the gradient
    // scatter sums runs of samples that share a grid cell)
    const uint32_t used_windows = std::min(windows, std::max(steps, 1u));
    const uint32_t buckets = kMlpSortCells * kMlpSortCells * kMlpSortCells;
    m->ordered.reserve(4 * (size_t)b * used_windows);
    m->sort_hist.reserve(buckets);
    m->sort_offs.reserve(buckets);
    m->sort_total.reserve(1);
    const MlpView v = mlp_view_of(m, false);
    for (uint32_t w = 0; w < used_windows; ++w)
      RTPG_HIP(launch_mlp_spatial_order(v, m->samples.p + 4 * (size_t)w * b, b, m->sort_hist.p, m->sort_offs.p,
                                        m->sort_total.p, m->ordered.p + 4 * (size_t)w * b, s));
    // every step's loss sum stays on the device until the last step
    m->loss.reserve(std::max(steps, 1u));
    for (uint32_t k = 0; k < steps; ++k) {
      const float4* smp = m->ordered.p + 4 * (size_t)((k % used_windows) * b);
      grad_pass(m, smp, b, k);
      adam_step(m, b);
    }
    m->count.zero(1, s);
    if (loss && steps > 0) {
      m->loss.download(loss, steps, s);
      for (uint32_t k = 0; k < steps; ++k) loss[k] /= (float)b;
    }
    RTPG_HIP(hipStreamSynchronize(s));
  });
}

int rtpg_mlp_export(rtpg_mlp m, float* weights, float* table) {
  return guarded([&] {
    require(m != nullptr, "NULL mlp");
    RTPG_HIP(hipSetDevice(m->ctx->device));
    if (weights) m->w.download(weights, kMlpWeights, m->ctx->stream);
    if (table) m->t.download(table, m->table_floats(), m->ctx->stream);
  });
}

int rtpg_mlp_cache_download(rtpg_mlp m, float* mix, float* c, uint32_t* entries) {
  return guarded([&] {
    require(m != nullptr && entries != nullptr, "NULL argument");
    require(m->cfg.cache_res > 0, "the mlp has no spatial cache (cache_res = 0)");
    RTPG_HIP(hipSetDevice(m->ctx->device));
    *entries = m->cache_entries();
    if (!mix && !c) return;
    mlp_cache_refresh(m);
    hipStream_t s = m->ctx->stream;
    if (mix) m->cache_mix.download(reinterpret_cast<float4*>(mix), (size_t)kMlpMixF4 * *entries, s);
    if (c) m->cache_c.download(c, *entries, s);
  });
}

int rtpg_mlp_import(rtpg_mlp m, const float* weights, const float* table) {
  return guarded([&] {
    require(m != nullptr, "NULL mlp");
    m->cache_stale = true;
    RTPG_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    if (weights) {
      std::vector<uint16_t> b(kMlpWeights);
      for (int k = 0; k < kMlpWeights; ++k) b[k] = f2bf(weights[k]);
      m->w.upload(weights, kMlpWeights, s);
      m->w_bf16.upload(b, s);
    }
    if (table) {
      size_t tf = m->table_floats();
      std::vector<uint32_t> b(tf / 2);
      for (size_t k = 0; k < tf / 2; ++k)
        b[k] = (uint32_t)f2bf(table[2 * k]) | ((uint32_t)f2bf(table[2 * k + 1]) << 16);
      m->t.upload(table, tf, s);
      m->t_bf16.upload(b, s);
    }
    RTPG_HIP(hipStreamSynchronize(s));
  });
}

int rtpg_mlp_info(rtpg_mlp m, uint64_t info[7], float res[16]) {
  return guarded([&] {
    require(m != nullptr, "NULL mlp");
    if (info) {
      uint32_t cnt = 0;
      m->count.download(&cnt, 1, m->ctx->stream);
      info[0] = kMlpWeights;
      info[1] = m->table_floats();
      info[2] = std::min(cnt, m->cfg.capacity);
      info[3] = m->steps;
      info[4] = kMlpLevels;
      info[5] = m->entries;
      info[6] = m->dense;
    }
    if (res)
      for (int l = 0; l < kMlpLevels; ++l) res[l] = m->res[l];
  });
}

}  // extern "C"
