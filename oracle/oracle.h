/*
 * oracle.h — CPU oracle for the rtpg path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the reference's algorithms on the hot path, used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product (rtlib_amd/, librtpg.so) never links or calls it.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the reference repository; Utils = RTLib/Src/Ext/OPX7/Include/RTLib/Ext/OPX7/
 * Utils, Math = RTLib/Src/Ext/CUDA/Include/RTLib/Ext/CUDA/Math, Test =
 * RTLib/Src/Ext/OPX7/Test).  How each is pinned (golden vectors, closed
 * forms, invariants, brute force) is stated in tests/ next to its test.
 */
#ifndef RTPG_ORACLE_H
#define RTPG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers ---- */
void orc_pcg32_raw(uint64_t initstate, uint64_t seq, uint32_t n, uint32_t* out);
void orc_rng_stream(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t n, uint32_t* out);
void orc_rng_floats(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t n, float* out);
float orc_to_uniform_float(uint32_t v);
uint32_t orc_xorshift32(uint32_t* state);
uint32_t orc_pcg1d(uint32_t v);
uint32_t orc_xxhash32(uint32_t data, uint32_t seed);

/* ---- Morton codes ---- */
uint32_t orc_morton2(uint32_t x, uint32_t y);
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z);

/* ---- canonical directions ---- */
void orc_dir_to_canonical(const float* d, float* uv);
void orc_canonical_to_dir(const float* uv, float* d);

/* ---- hash grid ---- */
typedef struct orc_grid {
  float origin[3], size[3];
  uint32_t bounds[3];
  uint32_t slots;
  uint32_t key_mode;
} orc_grid;
void orc_grid_cell_hash(const orc_grid* g, const float* pos, uint32_t* cell3, uint32_t* bucket, uint32_t* checksum);
/* insert = 0: previous-table probe; insert = 1: sequential insert into cur. */
void orc_grid_find(const orc_grid* g, uint32_t* cur, const uint32_t* prv, const float* pos, uint32_t n, int insert,
                   uint32_t* out);

/* ---- Morton quadtree ---- */
void orc_mq_record(float* w, uint32_t per_cell, uint32_t level, const uint32_t* slot, const float* canon,
                   const float* val, uint32_t n);
void orc_mq_build(float* w, uint32_t level, uint32_t per_cell, uint32_t cells);
void orc_mq_sample(const float* w, uint32_t per_cell, uint32_t level, const uint32_t* slot, uint32_t n,
                   uint64_t seed, float* out);
void orc_mq_pdf(const float* w, uint32_t per_cell, uint32_t level, const uint32_t* slot, const float* canon,
                uint32_t n, float* out);

/* ---- SD-tree ----
 * DTree node pools: sums[4*k], child[4*k] (u16, relative to the DTree base). */
void orc_dtree_build(uint32_t n_dtrees, const uint32_t* cnt, float* sums, const uint16_t* child, float* out_sum,
                     float* out_area);
/* Reset one building DTree from a sampling DTree (RTDTreeT::Reset).  Returns
 * the node count written to out_child (capacity 65536 nodes); *depth gets the
 * reached depth. */
uint32_t orc_dtree_reset(const float* src_sums, const uint16_t* src_child, float total, uint32_t max_depth,
                         float th, uint16_t* out_child, uint32_t* depth);
/* RTSTreeT::Refine on an encoded spatial tree (bit 31 = leaf | DTree index,
 * else first child).  Writes the refined tree and, per new DTree, the index of
 * the DTree it copies.  Returns the number of nodes, or 0 if cap is too small. */
uint32_t orc_stree_refine(const uint32_t* stree, uint32_t n_nodes, const float* b_sw, const uint32_t* b_cnt,
                          const uint32_t* s_cnt, uint64_t th, int max_mb, uint32_t* out, uint32_t cap,
                          uint32_t* map, uint32_t* n_dtrees_out, int octree);
uint64_t orc_split_threshold(int mode, float c, int iter, int spp);
void orc_stree_lookup(const uint32_t* stree, const float* aabb_min, float size, const float* pos, uint32_t n,
                      uint32_t* out);
void orc_sdtree_record(const uint32_t* stree, const float* aabb_min, float size, const uint32_t* b_base,
                       float* b_sw, float* b_sums, const uint16_t* b_child, const float* pos, const float* canon,
                       const float* irr, const float* sw, uint32_t n, uint32_t filter, uint64_t seed,
                       uint32_t batch);
void orc_sdtree_record_box(const uint32_t* stree, const float* aabb_min, float size, const uint32_t* b_base,
                           float* b_sw, float* b_sums, const uint16_t* b_child, const float* pos,
                           const float* canon, const float* irr, const float* sw, uint32_t n, int spatial,
                           int directional);
void orc_sdtree_sample(const uint32_t* s_base, const float* s_sum, const float* s_sw, const float* s_sums,
                       const uint16_t* s_child, const uint32_t* cells, uint32_t n, uint64_t seed, float* out);
void orc_sdtree_pdf(const uint32_t* s_base, const float* s_sum, const float* s_sw, const float* s_sums,
                    const uint16_t* s_child, const uint32_t* cells, const float* canon, uint32_t n, float* out);

/* ---- geometry ---- */
int orc_isect_tri(const float* o, const float* d, const float* v0, const float* e1, const float* e2, float tmin,
                  float* tbest, float* u, float* v);
/* Brute force over all triangles (verts/idx); hits: t, u, v, prim (bits). */
void orc_brute_closest(const float* verts, const uint32_t* idx, uint32_t nt, const float* rays, uint32_t n,
                       float* hits);

/* ---- integrator ---- */
typedef struct orc_scene orc_scene;
void orc_scene_set_light_pdf(orc_scene* s, const float* pdf);
orc_scene* orc_scene_create(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* albedo,
                            const uint32_t* mat, const float* emission, const float* spec);
void orc_scene_destroy(orc_scene* s);
void orc_scene_set_attributes(orc_scene* s, const float* texcoord, const uint32_t* tex_id, const float* vnormal,
                              const uint32_t* texels, uint64_t n_texels, const uint32_t* tex, uint32_t n_tex);
void orc_tex_sample(const uint32_t* texels, const uint32_t* tex, float u, float v, float* out3);
typedef struct orc_params {
  uint32_t width, height, flags, max_depth, rr_depth, num_candidates, spp, sample_offset;
  uint64_t seed;
  float bg[3];
  float fraction;
  float eye[3], u[3], v[3], w[3];
  uint32_t px_begin, px_end; /* pixel range [begin, end) to render */
  uint32_t threads;
} orc_params;
/* SD-tree guide view for the integrator (NULL for plain rendering): sampling
 * arrays are read; when record is set, building arrays are recorded into. */
typedef struct orc_sdtree {
  const uint32_t* stree;
  float aabb_min[3];
  float size;
  const uint32_t* s_base;
  const float* s_sum;
  const float* s_sw;
  const float* s_sums;
  const uint16_t* s_child;
  const uint32_t* b_base;
  float* b_sw;
  float* b_sums;
  const uint16_t* b_child;
  uint32_t filter; /* recording filters: 1 spatial box, 2 directional box (stochastic form) */
} orc_sdtree;
/* Hash-grid guide view for the integrator (SimpleGuide2.cu): the grid, its
 * current table (inserted into in a Locate pass) and previous table (probed),
 * and the Morton pyramids of level `level` (per_cell floats per slot):
 * sampling pyramids read, building pyramids recorded into. */
typedef struct orc_hashguide {
  orc_grid grid;
  uint32_t* cur;
  const uint32_t* prv;
  uint32_t level, per_cell;
  const float* s_w;
  float* b_w;
} orc_hashguide;
/* Debug frame of type 0-6 or 8 (raw quantity per pixel, 3 floats) and the
 * hit triangle per pixel (-1: miss). */
void orc_debug_frame(const orc_scene* s, const orc_params* p, uint32_t type, float* out3, int32_t* tri_out);
/* accum: 3*W*H floats, added to; returns the number of radiance rays traced. */
uint64_t orc_render(const orc_scene* s, const orc_params* p, const orc_sdtree* guide, float* accum);
/* The same integrator guided by the hash grid (RTPG flags USE_GRID, LOCATE,
 * GUIDE_SAMPLE, FINAL as in rtpg_render). */
uint64_t orc_render_grid(const orc_scene* s, const orc_params* p, const orc_hashguide* guide, float* accum);
/* FP64-accumulated gradient of sum_pixels dldi . image w.r.t. per-triangle
 * albedo (3*nt, Lambertian triangles) and per-light radiance (3*nl). */
void orc_adjoint(const orc_scene* s, const orc_params* p, const float* dldi, double* grad_alb, double* grad_le);

/* ---- pass controllers ----
 * RTSTreeControllerT (morton = 0) or RTMortonQuadTreeControllerT (morton = 1)
 * driven one launch of spl samples at a time, as the application drives them,
 * until the controller stops executing.  Per launch k: the state in force
 * (0 Record, 1 RecordAndSample, 2 Sample, 3 Locate), the iteration, the
 * samples of the current pass, and event bits: 1 the SD-tree is Reset before
 * the launch, 2 the launch ends a pass (SD-tree Build), 4 Morton Update and
 * pyramid build, 8 the application's hash-grid snapshot.  Returns the number
 * of launches (at most cap). */
uint32_t orc_ctrl_schedule(int morton, uint32_t budget, uint32_t iter_built, uint32_t iter_locate, float ratio,
                           uint32_t spl, uint32_t cap, uint32_t* state, uint32_t* iter, uint32_t* pass, uint32_t* ev);

#ifdef __cplusplus
}
#endif
#endif
