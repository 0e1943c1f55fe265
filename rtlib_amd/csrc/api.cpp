// C-ABI: context, scene upload, camera/frame, render and ray queries.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>

#include "../../include/rtpg.h"
#include "context.h"

namespace rtpg {

static thread_local std::string g_last_error;

void set_last_error(const std::string& s) { g_last_error = s; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw Error(RTPG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
}

template <class F>
static int guarded(F&& f) {
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

static void require(bool c, const char* msg) {
  if (!c) throw Error(RTPG_ERR_INVALID, msg);
}

}  // namespace rtpg

using namespace rtpg;

SceneView rtpg_context_t::scene_view() const {
  SceneView v{};
  v.nodes = nodes.p;
  v.tris = tris.p;
  v.shade = shade.p;
  v.spec = has_spec ? spec.p : nullptr;
  v.info = info.p;
  v.lights = lights.p;
  v.light_cdf = light_cdf.p;
  v.n_lights = n_lights;
  v.n_tris = n_tris;
  v.n_nodes = n_nodes;
  v.max_stack = max_stack;
  v.tattr = has_tattr ? tattr.p : nullptr;
  v.vnorm = has_vnorm ? vnorm.p : nullptr;
  v.tex = has_tattr ? tex.p : nullptr;
  v.texels = has_tattr ? texels.p : nullptr;
  return v;
}

extern "C" {

const char* rtpg_last_error(void) { return g_last_error.c_str(); }
const char* rtpg_version(void) { return "rtpg 0.1 (gfx950)"; }

int rtpg_context_create(int device, rtpg_context* out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    int count = 0;
    RTPG_HIP(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, "device index out of range");
    RTPG_HIP(hipSetDevice(device));
    scratch_setup(device);
    auto* c = new rtpg_context_t();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      hip_check(e, "hipStreamCreateWithFlags");
    }
    c->stream = c->own_stream;
    c->p_err.alloc(4);
    c->p_err.zero(4, c->stream);
    c->p_stats.alloc(4);
    *out = c;
  });
}

int rtpg_context_destroy(rtpg_context ctx) {
  return guarded([&] {
    if (!ctx) return;
    require(ctx->children == 0, "objects created on this context (SD-trees, hash grids, Morton quadtrees, "
                                "learned guides) are still alive: destroy them before the context");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    hipStream_t own = ctx->own_stream;
    delete ctx;
    if (own) (void)hipStreamDestroy(own);
  });
}

int rtpg_context_set_stream(rtpg_context ctx, void* hip_stream) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
  });
}

int rtpg_synchronize(rtpg_context ctx) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    RTPG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int rtpg_scene_upload(rtpg_context ctx, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* albedo, const uint32_t* mat_type, const float* emission, const float* spec,
                      uint32_t light_mode, const uint32_t* light_mesh) {
  return guarded([&] {
    require(ctx && verts && idx && albedo && mat_type && emission, "NULL argument");
    require(nt > 0 && nv > 0, "empty scene");
    require(light_mode <= 1u, "light_mode must be 0 or 1");
    require(light_mode == 0u || light_mesh != nullptr, "light_mode 1 needs light_mesh");
    RTPG_HIP(hipSetDevice(ctx->device));
    int threads = (int)std::max(1u, std::thread::hardware_concurrency());
    require(nt < (1u << 27), "at most 2^27 - 1 triangles (leaf entries hold a 27-bit first index)");
    BvhBuildResult b = build_bvh4(verts, nv, idx, nt, std::min(threads, 16));
    if (b.max_stack > 128u) throw Error(RTPG_ERR_LIMIT, "BVH too deep for the traversal stack");
    std::vector<float4> shade(nt), specv(nt);
    std::vector<uint32_t> info(nt);
    std::vector<uint32_t> slot_of(nt);
    for (uint32_t s = 0; s < nt; ++s) slot_of[b.order[s]] = s;
    // Light list in original triangle order.
    std::vector<uint32_t> light_tris;
    for (uint32_t t = 0; t < nt; ++t) {
      const float* e = emission + 3 * t;
      if (e[0] + e[1] + e[2] > 0.0f) {
        const TriRec& tr = b.tris[slot_of[t]];
        V3 e1{tr.e1[0], tr.e1[1], tr.e1[2]}, e2{tr.e2[0], tr.e2[1], tr.e2[2]};
        float area = 0.5f * length(cross(e1, e2));
        if (area > 0.0f) light_tris.push_back(t);
      }
    }
    uint32_t nl = (uint32_t)light_tris.size();
    std::vector<float> pdf(nl, 0.0f);
    if (light_mode == 0u) {
      for (uint32_t k = 0; k < nl; ++k) pdf[k] = 1.0f / (float)nl;
    } else {
      // Uniform mesh, then uniform triangle within the mesh.
      std::vector<uint32_t> meshes;
      for (uint32_t k = 0; k < nl; ++k) meshes.push_back(light_mesh[light_tris[k]]);
      std::vector<uint32_t> uniq = meshes;
      std::sort(uniq.begin(), uniq.end());
      uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
      for (uint32_t k = 0; k < nl; ++k) {
        uint32_t m = meshes[k];
        uint32_t cnt = (uint32_t)std::count(meshes.begin(), meshes.end(), m);
        pdf[k] = 1.0f / ((float)uniq.size() * (float)cnt);
      }
    }
    std::vector<float4> lights(4 * (size_t)std::max(nl, 1u));
    std::vector<float> cdf(std::max(nl, 1u), 1.0f);
    double acc = 0.0;
    for (uint32_t k = 0; k < nl; ++k) {
      uint32_t t = light_tris[k];
      uint32_t s = slot_of[t];
      const TriRec& tr = b.tris[s];
      V3 e1{tr.e1[0], tr.e1[1], tr.e1[2]}, e2{tr.e2[0], tr.e2[1], tr.e2[2]};
      float area = 0.5f * length(cross(e1, e2));
      lights[4 * k + 0] = make_float4(tr.v0[0], tr.v0[1], tr.v0[2], area);
      lights[4 * k + 1] = make_float4(e1.x, e1.y, e1.z, pdf[k]);
      lights[4 * k + 2] = make_float4(e2.x, e2.y, e2.z, bitsf(s));
      lights[4 * k + 3] = make_float4(emission[3 * t], emission[3 * t + 1], emission[3 * t + 2], 0.0f);
      acc += pdf[k];
      cdf[k] = (float)acc;
    }
    if (nl) cdf[nl - 1] = 1.0f;
    std::vector<int32_t> light_of(nt, -1);
    for (uint32_t k = 0; k < nl; ++k) light_of[light_tris[k]] = (int32_t)k;
    ctx->shade_host.assign(4 * (size_t)nt, 0.0f);
    for (uint32_t s = 0; s < nt; ++s) {
      uint32_t t = b.order[s];
      require(mat_type[t] <= 3u, "unknown material type");
      shade[s] = make_float4(albedo[4 * t], albedo[4 * t + 1], albedo[4 * t + 2], albedo[4 * t + 3]);
      for (int k = 0; k < 4; ++k) ctx->shade_host[4 * (size_t)s + k] = albedo[4 * t + k];
      specv[s] = spec ? make_float4(spec[4 * t], spec[4 * t + 1], spec[4 * t + 2], spec[4 * t + 3])
                      : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      uint32_t lw = light_of[t] >= 0 ? (uint32_t)light_of[t] + 1u : 0u;
      info[s] = (mat_type[t] & kTypeMask) | (lw << 4);
    }
    hipStream_t s = ctx->stream;
    ctx->nodes.upload(b.nodes, s);
    ctx->tris.upload(b.tris, s);
    ctx->shade.upload(shade, s);
    ctx->spec.upload(specv, s);
    ctx->info.upload(info, s);
    ctx->lights.upload(lights, s);
    ctx->light_cdf.upload(cdf, s);
    RTPG_HIP(hipStreamSynchronize(s));
    ctx->has_spec = spec != nullptr;
    ctx->has_tattr = ctx->has_vnorm = false;
    ctx->tattr.release();
    ctx->vnorm.release();
    ctx->tex.release();
    ctx->texels.release();
    ctx->lights_host = lights;
    ctx->light_tri = light_tris;
    ctx->order = b.order;
    ctx->slot_of = slot_of;
    ctx->n_tris = nt;
    ctx->n_lights = nl;
    ctx->n_nodes = (uint32_t)b.nodes.size();
    ctx->bvh_depth = b.max_depth;
    ctx->max_stack = b.max_stack;
    ctx->max_leaf = b.max_leaf;
    for (int k = 0; k < 3; ++k) {
      ctx->world[k] = b.world_lo[k];
      ctx->world[3 + k] = b.world_hi[k];
    }
  });
}

int rtpg_scene_attributes(rtpg_context ctx, const float* texcoord, const uint32_t* tex_id, const float* vnormal,
                          const uint32_t* texels, uint64_t n_texels, const uint32_t* tex, uint32_t n_tex) {
  return guarded([&] {
    require(ctx != nullptr, "NULL argument");
    require(ctx->n_tris > 0, "no scene");
    require(tex_id == nullptr || texcoord != nullptr, "tex_id needs texcoord");
    require(n_tex < kTexNone, "at most 65535 textures");
    require(n_tex == 0 || (tex != nullptr && texels != nullptr), "textures need tex and texels");
    for (uint32_t k = 0; k < n_tex; ++k) {
      const uint32_t* t = tex + 4 * (size_t)k;
      require(t[1] > 0 && t[2] > 0 && t[1] <= 32768u && t[2] <= 32768u, "texture size must be 1..32768");
      require(t[3] <= 1u, "texture address mode must be RTPG_TEX_CLAMP or RTPG_TEX_WRAP");
      require((uint64_t)t[0] + (uint64_t)t[1] * t[2] <= n_texels, "texture texels out of range");
    }
    const uint32_t nt = ctx->n_tris;
    RTPG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    bool any_tex = false;
    if (tex_id) {
      for (size_t k = 0; k < 4 * (size_t)nt; ++k) {
        require(tex_id[k] == RTPG_TEX_NONE || tex_id[k] < n_tex, "texture index out of range");
        any_tex |= tex_id[k] != RTPG_TEX_NONE;
      }
    }
    // leaf order; 16-bit texture indices, kTexNone = factor 1
    if (any_tex) {
      std::vector<float4> ta(2 * (size_t)nt);
      auto id16 = [&](size_t k) { return tex_id[k] == RTPG_TEX_NONE ? kTexNone : tex_id[k]; };
      for (uint32_t sl = 0; sl < nt; ++sl) {
        size_t t = ctx->order[sl];
        const float* c = texcoord + 6 * t;
        ta[2 * (size_t)sl] = make_float4(c[0], c[1], c[2], c[3]);
        ta[2 * (size_t)sl + 1] = make_float4(c[4], c[5], bitsf(id16(4 * t) | id16(4 * t + 1) << 16),
                                             bitsf(id16(4 * t + 2) | id16(4 * t + 3) << 16));
      }
      std::vector<uint4> td(n_tex);
      for (uint32_t k = 0; k < n_tex; ++k) td[k] = make_uint4(tex[4 * k], tex[4 * k + 1], tex[4 * k + 2], tex[4 * k + 3]);
      ctx->tattr.upload(ta, s);
      ctx->tex.upload(td, s);
      ctx->texels.upload(texels, (size_t)n_texels, s);
    } else {
      ctx->tattr.release();
      ctx->tex.release();
      ctx->texels.release();
    }
    if (vnormal) {
      std::vector<float4> vn(3 * (size_t)nt);
      for (uint32_t sl = 0; sl < nt; ++sl) {
        const float* q = vnormal + 9 * (size_t)ctx->order[sl];
        vn[3 * (size_t)sl] = make_float4(q[0], q[1], q[2], q[3]);
        vn[3 * (size_t)sl + 1] = make_float4(q[4], q[5], q[6], q[7]);
        vn[3 * (size_t)sl + 2] = make_float4(q[8], 0.0f, 0.0f, 0.0f);
      }
      ctx->vnorm.upload(vn, s);
    } else {
      ctx->vnorm.release();
    }
    RTPG_HIP(hipStreamSynchronize(s));
    ctx->has_tattr = any_tex;
    ctx->has_vnorm = vnormal != nullptr;
  });
}

int rtpg_scene_set_albedo(rtpg_context ctx, const float* albedo_rgb) {
  return guarded([&] {
    require(ctx && albedo_rgb, "NULL argument");
    require(ctx->n_tris > 0, "no scene");
    std::vector<float4> shade(ctx->n_tris);
    for (uint32_t s = 0; s < ctx->n_tris; ++s) {
      uint32_t t = ctx->order[s];
      float* h = &ctx->shade_host[4 * (size_t)s];
      h[0] = albedo_rgb[3 * t];
      h[1] = albedo_rgb[3 * t + 1];
      h[2] = albedo_rgb[3 * t + 2];
      shade[s] = make_float4(h[0], h[1], h[2], h[3]);
    }
    ctx->shade.upload(shade, ctx->stream);
    RTPG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int rtpg_scene_info(rtpg_context ctx, uint32_t info[6], float aabb[6]) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    if (info) {
      info[0] = ctx->n_nodes;
      info[1] = ctx->bvh_depth;
      info[2] = ctx->max_stack;
      info[3] = ctx->max_leaf;
      info[4] = ctx->n_lights;
      info[5] = ctx->n_tris;
    }
    if (aabb) std::memcpy(aabb, ctx->world, sizeof(ctx->world));
  });
}

int rtpg_camera_set(rtpg_context ctx, const float eye[3], const float u[3], const float v[3], const float w[3]) {
  return guarded([&] {
    require(ctx && eye && u && v && w, "NULL argument");
    for (int k = 0; k < 3; ++k) {
      ctx->cam.eye[k] = eye[k];
      ctx->cam.u[k] = u[k];
      ctx->cam.v[k] = v[k];
      ctx->cam.w[k] = w[k];
    }
  });
}

int rtpg_frame_resize(rtpg_context ctx, uint32_t width, uint32_t height) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    require(width > 0 && height > 0 && (uint64_t)width * height < (1ull << 31), "bad frame size");
    ctx->width = width;
    ctx->height = height;
    ctx->accum.alloc(3 * (size_t)width * height);
    ctx->accum.zero(3 * (size_t)width * height, ctx->stream);
    ctx->accum_spp = 0;
  });
}

int rtpg_accum_clear(rtpg_context ctx) {
  return guarded([&] {
    require(ctx && ctx->accum.p, "frame not sized");
    ctx->accum.zero(3 * (size_t)ctx->width * ctx->height, ctx->stream);
    ctx->accum_spp = 0;
  });
}

int rtpg_accum_download(rtpg_context ctx, float* out, uint32_t* spp) {
  return guarded([&] {
    require(ctx && ctx->accum.p, "frame not sized");
    if (out) ctx->accum.download(out, 3 * (size_t)ctx->width * ctx->height, ctx->stream);
    if (spp) *spp = ctx->accum_spp;
  });
}

int rtpg_frame_tonemap(rtpg_context ctx, uint8_t* rgba) {
  return guarded([&] {
    require(ctx && ctx->accum.p && rgba, "NULL argument or frame not sized");
    RTPG_HIP(hipSetDevice(ctx->device));
    uint32_t npix = ctx->width * ctx->height;
    DBuf<uchar4> frame;
    frame.alloc(npix);
    float inv = ctx->accum_spp ? 1.0f / (float)ctx->accum_spp : 0.0f;
    RTPG_HIP(launch_tonemap(ctx->accum.p, npix, inv, frame.p, ctx->stream));
    frame.download(reinterpret_cast<uchar4*>(rgba), npix, ctx->stream);
  });
}

}  // extern "C"

namespace rtpg {
__global__ void k_stats_add(const uint32_t* count, const uint32_t* scount, uint32_t depth, uint32_t n,
                            unsigned long long* stats);
hipError_t launch_stats_add(const uint32_t* count, const uint32_t* scount, uint32_t depth, uint32_t n,
                            unsigned long long* stats, hipStream_t s);
SdTreeView sdtree_view_of(rtpg_sdtree_t* t);
void guide_views(const rtpg_render_params* rp, rtpg_context_t* ctx, HashGridView& hg, MortonView& mq);
}  // namespace rtpg

static void ensure_arena(rtpg_context_t* c, uint32_t n, uint32_t max_depth, bool guide, uint32_t mlp_rec,
                         uint32_t mix_slots) {
  c->p_ray_o.reserve(n);
  c->p_ray_d.reserve(n);
  c->p_T.reserve(n);
  c->p_rng.reserve(n);
  c->p_ray_o1.reserve(n);
  c->p_ray_d1.reserve(n);
  c->p_T1.reserve(n);
  c->p_rng1.reserve(n);
  c->p_L.reserve(n);
  c->p_hit.reserve(n);
  c->p_nhit.reserve(n);
  c->p_q0.reserve(n);
  c->p_q1.reserve(n);
  c->p_shadow.reserve(3 * (size_t)n);
  c->p_count.reserve(2 * (size_t)(max_depth + 2) * kQueueRegions);
  // spill stack: (entry, distance) pairs beyond the kShortStack LDS entries
  if (c->max_stack > kShortStack) c->p_deep.reserve(2 * (size_t)(c->max_stack - kShortStack) * n);
  if (guide) {
    c->p_vA.reserve((size_t)max_depth * n);
    c->p_vB.reserve((size_t)max_depth * n);
    c->p_vC.reserve((size_t)max_depth * n);
  }
  if (mix_slots) {
    c->p_mix.reserve((size_t)kMlpMixF4 * mix_slots);
    c->p_mixc.reserve(mix_slots);
  }
  if (mlp_rec) c->p_mv.reserve((size_t)5 * max_depth * mlp_rec);
}

static PathSoA make_soa(rtpg_context_t* ctx, uint32_t n, uint32_t max_depth, bool guide, bool counting,
                        rtpg_mlp_t* mlp = nullptr) {
  // learned guide: every stride-th path records (at most record_paths slots)
  uint32_t mlp_stride = mlp ? std::max(1u, (uint32_t)(((uint64_t)n + mlp->cfg.record_paths - 1u) / mlp->cfg.record_paths)) : 0u;
  uint32_t mlp_rec = mlp ? (n + mlp_stride - 1u) / mlp_stride : 0u;
  // per-slot mixtures of the inference pass (none with the spatial cache)
  uint32_t mix_slots = (mlp && mlp->cfg.cache_res == 0) ? n : 0u;
  ensure_arena(ctx, n, max_depth, guide, mlp_rec, mix_slots);
  PathSoA ps{};
  ps.n = n;
  ps.ray_o[0] = ctx->p_ray_o.p;
  ps.ray_o[1] = ctx->p_ray_o1.p;
  ps.ray_d[0] = ctx->p_ray_d.p;
  ps.ray_d[1] = ctx->p_ray_d1.p;
  ps.T[0] = ctx->p_T.p;
  ps.T[1] = ctx->p_T1.p;
  ps.rng[0] = ctx->p_rng.p;
  ps.rng[1] = ctx->p_rng1.p;
  ps.L = ctx->p_L.p;
  ps.hit = ctx->p_hit.p;
  ps.nhit = ctx->p_nhit.p;
  ps.queue[0] = ctx->p_q0.p;
  ps.queue[1] = ctx->p_q1.p;
  // queue regions: slots per region a multiple of 1024 (every work-group of a
  // queue kernel lies in one region), at most kQueueRegions of them
  const uint32_t per = (n + kQueueRegions - 1u) / kQueueRegions;
  ps.qreg = std::max(1024u, (per + 1023u) & ~1023u);
  ps.nreg = (n + ps.qreg - 1u) / ps.qreg;
  const uint32_t ncnt = (max_depth + 2u) * kQueueRegions;
  ps.count = ctx->p_count.p;
  ps.scount = ctx->p_count.p + ncnt;
  ps.shadow = ctx->p_shadow.p;
  ps.deep = ctx->max_stack > kShortStack ? ctx->p_deep.p : nullptr;
  ps.err = ctx->p_err.p;
  ps.vA = guide ? ctx->p_vA.p : nullptr;
  ps.vB = guide ? ctx->p_vB.p : nullptr;
  ps.vC = guide ? ctx->p_vC.p : nullptr;
  ps.work = counting ? ctx->p_work.p : nullptr;
  ps.npix = ctx->width * ctx->height;
  ps.rec_limit = n;
  if (mlp) {
    ps.mix = ctx->p_mix.p;
    ps.mixc = ctx->p_mixc.p;
    ps.mv = ctx->p_mv.p;
    ps.mlp_rec = mlp_rec;
    ps.mlp_stride = mlp_stride;
    ps.mlp_phase = 0;
    ps.mlp_guide_depth = mlp->cfg.guide_depth ? mlp->cfg.guide_depth : 0xFFFFFFFFu;
    ps.mlp_out = mlp->samples.p;
    ps.mlp_count = mlp->count.p;
    ps.mlp_cap = mlp->cfg.capacity;
    if (mlp->cfg.cache_res > 0) {
      ps.mcache = mlp->cache_mix.p;
      ps.mcache_c = mlp->cache_c.p;
      ps.mcache_res = mlp->cfg.cache_res;
      for (int k = 0; k < 3; ++k) ps.mcache_lo[k] = mlp->lo[k];
      ps.mcache_inv = 1.0f / mlp->size;
    }
  }
  return ps;
}

static RenderParams make_rp(rtpg_context_t* ctx, const rtpg_render_params* prm, uint32_t spp, uint32_t done) {
  RenderParams rp{};
  rp.width = ctx->width;
  rp.height = ctx->height;
  rp.flags = prm->flags;
  rp.max_depth = prm->max_depth;
  rp.rr_depth = prm->rr_depth;
  rp.num_candidates = prm->num_candidates;
  rp.spp = spp;
  rp.sample_offset = prm->sample_offset + done;
  rp.seed = prm->seed;
  rp.bg[0] = prm->bg[0];
  rp.bg[1] = prm->bg[1];
  rp.bg[2] = prm->bg[2];
  rp.fraction = prm->fraction;
  return rp;
}

extern "C" int rtpg_adjoint(rtpg_context ctx, const rtpg_render_params* prm, const float* dl_dimage,
                            float* grad_albedo, float* grad_emission) {
  return guarded([&] {
    require(ctx && prm && dl_dimage, "NULL argument");
    require(ctx->n_tris > 0 && ctx->accum.p != nullptr, "no scene or frame");
    require(prm->max_depth >= 1 && prm->max_depth <= 64, "max_depth must be in [1, 64]");
    require(prm->spp >= 1, "spp must be >= 1");
    require(!(prm->flags & (RTPG_FLAG_USE_TREE | RTPG_FLAG_USE_GRID | RTPG_FLAG_USE_MLP)),
            "the adjoint renders without guides");
    require(!(prm->flags & RTPG_FLAG_RIS), "the adjoint does not support RIS (its resampling depends on albedo)");
    require(!(prm->flags & RTPG_FLAG_RR), "the adjoint needs Russian roulette off (its survival depends on albedo)");
    RTPG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    ctx->p_err.zero(1, s);  // the stack-bound flag reports this call only
    uint32_t npix = ctx->width * ctx->height;
    uint32_t nl = std::max(ctx->n_lights, 1u);
    // buffers kept by the context between calls (they only grow)
    DBuf<float>& d_dldi = ctx->adj_dldi;
    DBuf<float>& d_galb = ctx->adj_galb;
    DBuf<float>& d_gle = ctx->adj_gle;
    DBuf<float>& scratch = ctx->adj_scratch;
    DBuf<float4>& d_ltot = ctx->adj_ltot;
    d_dldi.upload(dl_dimage, 3 * (size_t)npix, s);
    d_galb.reserve(3 * (size_t)ctx->n_tris);
    d_galb.zero(3 * (size_t)ctx->n_tris, s);
    d_gle.reserve(3 * (size_t)nl);
    d_gle.zero(3 * (size_t)nl, s);
    scratch.reserve(3 * (size_t)npix);
    uint32_t max_paths = std::min(prm->max_paths ? prm->max_paths : (1u << 25), 1u << 26);
    uint32_t chunk_spp = std::min(std::max(1u, max_paths / npix), prm->spp);
    SceneView sc = ctx->scene_view();
    SdTreeView sd{};
    HashGridView hg{};
    MortonView mq{};
    uint32_t done = 0;
    while (done < prm->spp) {
      uint32_t cs = std::min(chunk_spp, prm->spp - done);
      uint32_t n = cs * npix;
      PathSoA ps = make_soa(ctx, n, prm->max_depth, false, false);
      d_ltot.reserve(n);
      ps.dldi = d_dldi.p;
      ps.adjL = d_ltot.p;
      ps.grad_alb = d_galb.p;
      ps.grad_le = d_gle.p;
      ps.inv_spp = 1.0f / (float)prm->spp;
      RenderParams rp = make_rp(ctx, prm, cs, done);
      // primal pass: the total radiance of every path
      ps.adj_mode = 0u;
      RTPG_HIP(launch_render_chunk(rp, ctx->cam, sc, ps, sd, hg, mq, MlpView{}, 0, scratch.p, s));
      RTPG_HIP(hipMemcpyAsync(d_ltot.p, ps.L, sizeof(float4) * n, hipMemcpyDeviceToDevice, s));
      // replay with the same random streams; scatter the derivatives
      ps.adj_mode = 1u;
      RTPG_HIP(launch_render_chunk(rp, ctx->cam, sc, ps, sd, hg, mq, MlpView{}, 0, scratch.p, s));
      done += cs;
    }
    std::vector<float>& galb = ctx->adj_galb_host;
    galb.resize(3 * (size_t)ctx->n_tris);
    d_galb.download(galb.data(), galb.size(), s);
    std::vector<float> gle = d_gle.to_vector(3 * (size_t)nl, s);
    uint32_t err = 0;
    ctx->p_err.download(&err, 1, s);
    if (err) throw Error(RTPG_ERR_LIMIT, "traversal stack bound exceeded");
    if (grad_albedo) {
      for (uint32_t slot = 0; slot < ctx->n_tris; ++slot) {
        uint32_t t = ctx->order[slot];
        for (int k = 0; k < 3; ++k) grad_albedo[3 * (size_t)t + k] += galb[3 * (size_t)slot + k];
      }
    }
    if (grad_emission) {
      for (uint32_t l = 0; l < ctx->n_lights; ++l)
        for (int k = 0; k < 3; ++k) grad_emission[3 * (size_t)l + k] += gle[3 * (size_t)l + k];
    }
  });
}

extern "C" int rtpg_scene_set_emission(rtpg_context ctx, const float* le_rgb) {
  return guarded([&] {
    require(ctx && le_rgb, "NULL argument");
    require(ctx->n_lights > 0, "scene has no lights");
    for (uint32_t l = 0; l < ctx->n_lights; ++l) {
      ctx->lights_host[4 * (size_t)l + 3] = make_float4(le_rgb[3 * l], le_rgb[3 * l + 1], le_rgb[3 * l + 2], 0.0f);
    }
    ctx->lights.upload(ctx->lights_host, ctx->stream);
    RTPG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int rtpg_scene_lights(rtpg_context ctx, uint32_t* light_tri) {
  return guarded([&] {
    require(ctx && light_tri, "NULL argument");
    for (uint32_t l = 0; l < ctx->n_lights; ++l) light_tri[l] = ctx->light_tri[l];
  });
}

extern "C" int rtpg_render(rtpg_context ctx, const rtpg_render_params* prm) {
  return guarded([&] {
    require(ctx && prm, "NULL argument");
    require(ctx->n_tris > 0, "no scene uploaded");
    require(ctx->accum.p != nullptr, "frame not sized");
    require(prm->max_depth >= 1 && prm->max_depth <= 64, "max_depth must be in [1, 64]");
    require(!(prm->flags & RTPG_FLAG_RIS) || prm->num_candidates >= 1, "RIS needs num_candidates >= 1");
    RTPG_HIP(hipSetDevice(ctx->device));
    const bool use_tree = (prm->flags & RTPG_FLAG_USE_TREE) != 0;
    const bool use_grid = (prm->flags & RTPG_FLAG_USE_GRID) != 0;
    const bool use_mlp = (prm->flags & RTPG_FLAG_USE_MLP) != 0;
    require((int)use_tree + (int)use_grid + (int)use_mlp <= 1, "USE_TREE, USE_GRID and USE_MLP are exclusive");
    require(!use_tree || prm->sdtree != nullptr, "USE_TREE needs an sdtree");
    require(!use_grid || (prm->hashgrid != nullptr && prm->mqtree != nullptr), "USE_GRID needs hashgrid + mqtree");
    require(!use_mlp || prm->mlp != nullptr, "USE_MLP needs an mlp");
    require(!use_mlp || prm->mlp->ctx == ctx, "the mlp belongs to another context");
    int guide = use_tree ? 1 : (use_grid ? 2 : (use_mlp ? 3 : 0));
    require(guide == 0 || (prm->flags & RTPG_FLAG_FINAL) || prm->max_depth <= 16,
            "recording guide vertices supports max_depth <= 16");
    uint32_t npix = ctx->width * ctx->height;
    // paths per chunk: 32M by default (about 30 GB of path state when guiding);
    // at most 64M so that per-chunk record indices stay within 32 bits
    uint32_t max_paths = std::min(prm->max_paths ? prm->max_paths : (1u << 25), 1u << 26);
    uint32_t chunk_spp = std::max(1u, max_paths / npix);
    chunk_spp = std::min(chunk_spp, std::max(prm->spp, 1u));
    SceneView sc = ctx->scene_view();
    SdTreeView sd{};
    HashGridView hg{};
    MortonView mq{};
    MlpView ml{};
    if (guide == 1) sd = sdtree_view_of(prm->sdtree);
    // a learned-guide render that records also records into params->sdtree
    if (guide == 3 && prm->sdtree && !(prm->flags & RTPG_FLAG_FINAL)) sd = sdtree_view_of(prm->sdtree);
    if (guide == 2) guide_views(prm, ctx, hg, mq);
    if (guide == 3) {
      if (prm->flags & RTPG_FLAG_GUIDE_SAMPLE) mlp_cache_refresh(prm->mlp);
      ml = mlp_view_of(prm->mlp, false);
    }
    hipStream_t s = ctx->stream;
    // the stack-bound flag reports this call only
    ctx->p_err.zero(1, s);
    ctx->p_stats.zero(4, s);
    const bool moments = (prm->flags & RTPG_FLAG_MOMENTS) != 0;
    if (moments) {
      ctx->moments.reserve(2 * (size_t)npix);
      ctx->moments.zero(2 * (size_t)npix, s);
      ctx->moments_spp = prm->spp;
    }
    const bool timing = (ctx->prof_mode & 1u) != 0;
    const bool counting = (ctx->prof_mode & 2u) != 0;
    // events of up to kEvChunks chunks are read back together, so timing adds
    // one host synchronisation per kEvChunks chunks
    constexpr uint32_t kEvChunks = 8;
    const size_t ev_stride = (size_t)kTimedKernels * kMaxDepthEv * 2;
    uint32_t ev_pending = 0;
    auto drain_events = [&] {
      if (ev_pending == 0) return;
      RTPG_HIP(hipStreamSynchronize(s));
      for (uint32_t c = 0; c < ev_pending; ++c) {
        const hipEvent_t* ev = ctx->events.data() + c * ev_stride;
        for (int k = 0; k < kTimedKernels; ++k) {
          uint32_t nd = (k <= 2) ? prm->max_depth : 1u;
          for (uint32_t d = 0; d < nd; ++d) {
            float ms = 0.0f;
            RTPG_HIP(hipEventElapsedTime(&ms, ev[((size_t)k * kMaxDepthEv + d) * 2],
                                         ev[((size_t)k * kMaxDepthEv + d) * 2 + 1]));
            ctx->kernel_ms[k] += ms;
          }
        }
      }
      ev_pending = 0;
    };
    if (timing && ctx->events.empty()) {
      ctx->events.resize(kEvChunks * ev_stride);
      // timing-only events: no system-scope fence (cache writeback and
      // invalidate) at each record, which would slow the kernels being timed
      for (hipEvent_t& e : ctx->events) RTPG_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    }
    for (double& m : ctx->kernel_ms) m = 0.0;
    for (uint64_t& l : ctx->kernel_launches) l = 0;
    if (counting) {
      ctx->p_work.reserve(4);
      ctx->p_work.zero(4, s);
    }
    uint32_t done = 0;
    while (done < prm->spp) {
      uint32_t cs = std::min(chunk_spp, prm->spp - done);
      uint32_t n = cs * npix;
      PathSoA ps = make_soa(ctx, n, prm->max_depth, guide != 0, counting, guide == 3 ? prm->mlp : nullptr);
      RenderParams rp = make_rp(ctx, prm, cs, done);
      if (guide == 3) ps.mlp_phase = rp.sample_offset % ps.mlp_stride;
      ps.moments = moments ? ctx->moments.p : nullptr;
      if (timing && ev_pending == kEvChunks) drain_events();
      RTPG_HIP(launch_render_chunk(rp, ctx->cam, sc, ps, sd, hg, mq, ml, guide, ctx->accum.p, s,
                                   timing ? ctx->events.data() + (ev_pending++) * ev_stride : nullptr));
      RTPG_HIP(launch_stats_add(ps.count, ps.scount, prm->max_depth, n, ctx->p_stats.p, s));
      for (int k = 0; k < 3; ++k) ctx->kernel_launches[k] += prm->max_depth;
      ctx->kernel_launches[3] += 1;
      if (guide != 0 && !(prm->flags & RTPG_FLAG_FINAL)) ctx->kernel_launches[4] += 1;
      if (guide == 3 && (prm->flags & RTPG_FLAG_GUIDE_SAMPLE) && prm->mlp->cfg.cache_res == 0)
        ctx->kernel_launches[1] += prm->max_depth;
      done += cs;
    }
    if (timing) drain_events();
    if (counting) {
      unsigned long long w[4];
      ctx->p_work.download(w, 4, s);
      for (int k = 0; k < 4; ++k) ctx->work[k] = w[k];
    }
    ctx->accum_spp += prm->spp;
    unsigned long long st[4];
    ctx->p_stats.download(st, 4, s);
    uint32_t err = 0;
    ctx->p_err.download(&err, 1, s);
    if (err) throw Error(RTPG_ERR_LIMIT, "traversal or guide stack bound exceeded");
    ctx->last_stats[0] = st[0];
    ctx->last_stats[1] = st[1];
    ctx->last_stats[2] = st[2];
  });
}

extern "C" int rtpg_render_relvar(rtpg_context ctx, double* relvar) {
  return guarded([&] {
    require(ctx && relvar, "NULL argument");
    require(ctx->moments.p != nullptr && ctx->moments_spp >= 2,
            "needs a render with RTPG_FLAG_MOMENTS and at least 2 samples per pixel");
    RTPG_HIP(hipSetDevice(ctx->device));
    const uint32_t npix = ctx->width * ctx->height;
    const uint32_t blocks = (npix + 255u) / 256u;
    ctx->relvar_part.reserve(blocks);
    RTPG_HIP(launch_relvar(ctx->moments.p, ctx->accum.p, npix, ctx->moments_spp, ctx->accum_spp, ctx->relvar_part.p,
                           ctx->stream));
    std::vector<double> ps = ctx->relvar_part.to_vector(blocks, ctx->stream);
    double acc = 0.0;
    for (double v : ps) acc += v;
    *relvar = acc / (double)npix;
  });
}

extern "C" int rtpg_render_moments(rtpg_context ctx, float* sums, uint32_t* spp) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    require(ctx->moments.p != nullptr, "no render with RTPG_FLAG_MOMENTS yet");
    RTPG_HIP(hipSetDevice(ctx->device));
    if (spp) *spp = ctx->moments_spp;
    if (sums) ctx->moments.download(sums, 2 * (size_t)ctx->width * ctx->height, ctx->stream);
  });
}

extern "C" int rtpg_render_profile(rtpg_context ctx, uint32_t mode) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    require(mode <= 3u, "mode must be in [0, 3]");
    ctx->prof_mode = mode;
  });
}

extern "C" int rtpg_render_kernel_stats(rtpg_context ctx, double ms[5], uint64_t work[4], uint64_t launches[5]) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    if (ms) for (int k = 0; k < kTimedKernels; ++k) ms[k] = ctx->kernel_ms[k];
    if (work) for (int k = 0; k < 4; ++k) work[k] = ctx->work[k];
    if (launches) for (int k = 0; k < kTimedKernels; ++k) launches[k] = ctx->kernel_launches[k];
  });
}

extern "C" int rtpg_render_stats(rtpg_context ctx, uint64_t stats[3]) {
  return guarded([&] {
    require(ctx && stats, "NULL argument");
    for (int k = 0; k < 3; ++k) stats[k] = ctx->last_stats[k];
  });
}

static int trace_common(rtpg_context ctx, const float* rays, uint32_t n, float* hits, uint32_t* occl, bool any) {
  return guarded([&] {
    require(ctx && rays && (hits || occl), "NULL argument");
    require(ctx->n_tris > 0, "no scene uploaded");
    if (n == 0) return;
    RTPG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DBuf<float4> d_rays, d_hits;
    DBuf<uint32_t> d_occ, d_deep;
    d_rays.upload(reinterpret_cast<const float4*>(rays), 2 * (size_t)n, s);
    if (any) d_occ.alloc(n); else d_hits.alloc(n);
    if (ctx->max_stack > kShortStack) d_deep.alloc(2 * (size_t)(ctx->max_stack - kShortStack) * n);
    ctx->p_err.zero(1, s);  // the stack-bound flag reports this call only
    RTPG_HIP(launch_trace_rays(ctx->scene_view(), d_rays.p, n, d_hits.p, d_occ.p, any,
                               ctx->max_stack > kShortStack ? d_deep.p : nullptr, ctx->p_err.p, s));
    if (any) d_occ.download(occl, n, s); else d_hits.download(reinterpret_cast<float4*>(hits), n, s);
    uint32_t err = 0;
    ctx->p_err.download(&err, 1, s);
    if (err) throw Error(RTPG_ERR_LIMIT, "traversal stack bound exceeded");
  });
}

extern "C" int rtpg_render_debug(rtpg_context ctx, const rtpg_render_params* prm, uint32_t frame_type,
                                 float* value, uint8_t* rgba) {
  return guarded([&] {
    require(ctx && prm && (value || rgba), "NULL argument");
    require(ctx->n_tris > 0 && ctx->accum.p != nullptr, "no scene or frame");
    require(frame_type <= 8u, "frame_type must be in [0, 8]");
    const bool use_grid = frame_type == 7u && (prm->flags & RTPG_FLAG_USE_GRID) != 0;
    require(!use_grid || (prm->hashgrid != nullptr && prm->mqtree != nullptr), "USE_GRID needs hashgrid + mqtree");
    RTPG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    uint32_t npix = ctx->width * ctx->height;
    HashGridView hg{};
    MortonView mq{};
    if (use_grid) guide_views(prm, ctx, hg, mq);
    DBuf<float4> d_val;
    DBuf<uchar4> d_frame;
    DBuf<uint32_t> d_deep;
    if (value) d_val.alloc(npix);
    if (rgba) d_frame.alloc(npix);
    if (ctx->max_stack > kShortStack) d_deep.alloc(2 * (size_t)(ctx->max_stack - kShortStack) * npix);
    ctx->p_err.zero(1, s);
    RTPG_HIP(launch_debug(make_rp(ctx, prm, 1, 0), ctx->cam, ctx->scene_view(), hg, mq, frame_type, use_grid,
                          d_val.p, d_frame.p, ctx->max_stack > kShortStack ? d_deep.p : nullptr, ctx->p_err.p, s));
    if (value) d_val.download(reinterpret_cast<float4*>(value), npix, s);
    if (rgba) d_frame.download(reinterpret_cast<uchar4*>(rgba), npix, s);
    uint32_t err = 0;
    ctx->p_err.download(&err, 1, s);
    if (err) throw Error(RTPG_ERR_LIMIT, "traversal stack bound exceeded");
  });
}

extern "C" int rtpg_trace_closest(rtpg_context ctx, const float* rays, uint32_t n, float* hits) {
  return trace_common(ctx, rays, n, hits, nullptr, false);
}

extern "C" int rtpg_trace_occluded(rtpg_context ctx, const float* rays, uint32_t n, uint32_t* occluded) {
  return trace_common(ctx, rays, n, nullptr, occluded, true);
}
