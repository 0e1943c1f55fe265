/*
 * rtpg.h — C ABI of the MI355X (gfx950) path-tracing and path-guiding core.
 *
 * Drop-in boundary for the tracer/integrator path of RTLib's OptiX test
 * application.  The reference has no C ABI: its surface is C++ classes and
 * templates.  Each entry point below names the reference interface it replaces
 * (paths relative to the reference repository root; OPX7 = RTLib/Src/Ext/OPX7,
 * Test = OPX7/Test, Utils = OPX7/Include/RTLib/Ext/OPX7/Utils).
 *
 * Conventions
 *   - Every function returns 0 (RTPG_OK) on success or a negative status; no
 *     C++ exception crosses the ABI.  rtpg_last_error() returns a thread-local
 *     description of the last failure.
 *   - Host arrays are owned by the caller and only read/written during the call.
 *     Device memory is owned by the library.
 *   - All work is ordered on the context's HIP stream (set with
 *     rtpg_context_set_stream, default: the context's own stream).  Download
 *     functions synchronise the stream.  One host thread per context.
 *   - Handles are destroyed with their *_destroy function; destroying a context
 *     does not destroy guides created on it (destroy those first).
 */
#ifndef RTPG_H
#define RTPG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTPG_OK 0
#define RTPG_ERR_INVALID (-1)
#define RTPG_ERR_HIP (-2)
#define RTPG_ERR_STATE (-3)
#define RTPG_ERR_LIMIT (-4)

/* Render flags.  Bits 1..64 keep the meaning of the reference's ParamFlag
 * (Test/cuda/SimpleKernel.h:149-159). */
#define RTPG_FLAG_NEE 1u
#define RTPG_FLAG_RIS 2u
#define RTPG_FLAG_USE_GRID 4u
#define RTPG_FLAG_USE_TREE 8u
#define RTPG_FLAG_GUIDE_SAMPLE 16u /* reference PARAM_FLAG_BUILD: sample from the guide */
#define RTPG_FLAG_FINAL 32u        /* no recording */
#define RTPG_FLAG_LOCATE 64u       /* insert hash-grid cells */
#define RTPG_FLAG_MIS 128u
#define RTPG_FLAG_RR 256u
#define RTPG_FLAG_USE_MLP 512u
#define RTPG_FLAG_MOMENTS 1024u /* per-pixel moments of this render's samples (rtpg_render_relvar) */

/* Material types (per triangle). */
#define RTPG_MAT_LAMBERT 0u
#define RTPG_MAT_GGX 1u   /* param = roughness (alpha) */
#define RTPG_MAT_PHONG 2u /* reference "PHONG": param ignored, specular + shininess from spec[] */
#define RTPG_MAT_GLASS 3u /* param = index of refraction */

/* rtpg_scene_attributes: "no texture" index; texture address modes */
#define RTPG_TEX_NONE 0xFFFFFFFFu
#define RTPG_TEX_CLAMP 0u
#define RTPG_TEX_WRAP 1u

typedef struct rtpg_context_t* rtpg_context;
typedef struct rtpg_sdtree_t* rtpg_sdtree;
typedef struct rtpg_sdtree_ctrl_t* rtpg_sdtree_ctrl;
typedef struct rtpg_hashgrid_t* rtpg_hashgrid;
typedef struct rtpg_mqtree_t* rtpg_mqtree;
typedef struct rtpg_mqtree_ctrl_t* rtpg_mqtree_ctrl;
typedef struct rtpg_mlp_t* rtpg_mlp;

const char* rtpg_last_error(void);
const char* rtpg_version(void);

/* --- context ---------------------------------------------------------------
 * Replaces OPX7Context/CUDAContext creation (OPX7/Src/OPX7Context.cpp:274-292,
 * Ext/CUDA/Src/CUDAContext.cpp) and CUDAStream::Synchronize. */
int rtpg_context_create(int device, rtpg_context* out);
/* Fails (RTPG_ERR_INVALID, the context untouched) while SD-trees, hash grids,
 * Morton quadtrees or learned guides created on it are alive: they use its
 * device and stream, so destroy them first (the reference requires the same
 * order of its Destroy() calls). */
int rtpg_context_destroy(rtpg_context ctx);
int rtpg_context_set_stream(rtpg_context ctx, void* hip_stream);
int rtpg_synchronize(rtpg_context ctx);

/* --- scene ------------------------------------------------------------------
 * Replaces geometry upload + GAS/IAS build (Test/Include/RTLibExtOPX7Test.h:
 * 153-296, 515-656; OPX7/Src/OPX7Natives.cpp:195-244), the hitgroup material
 * records (RTLibExtOPX7Test.h:1602-1638) and the mesh-light list upload
 * (RTLibExtOPX7Test.h:700-725).
 *   verts:    3*nv floats;  idx: 3*nt vertex indices
 *   albedo:   4*nt floats (rgb, param) per triangle
 *   mat_type: nt values (RTPG_MAT_*)
 *   emission: 3*nt floats (0 = not emissive; emissive triangles form the light list)
 *   spec:     4*nt floats (specular rgb, shininess) or NULL
 * light_mode 0: lights picked uniformly per triangle; 1: pick a light mesh
 * uniformly, then a triangle in it (reference MeshLightList::Sample,
 * SimpleKernel.h:81-120), with light_mesh[nt] giving each triangle's mesh id. */
int rtpg_scene_upload(rtpg_context ctx, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* albedo, const uint32_t* mat_type, const float* emission, const float* spec,
                      uint32_t light_mode, const uint32_t* light_mesh);
/* Replace per-triangle albedo (rgb, 3*nt) in original triangle order. */
int rtpg_scene_set_albedo(rtpg_context ctx, const float* albedo_rgb);
/* Texture maps and per-vertex attributes of the uploaded triangles, in the
 * triangle order of rtpg_scene_upload.  Replaces TextureData::LoadFromPath and
 * the hitgroup/mesh-light texture objects (RTLibExtOPX7Test.h:386-511,
 * 1602-1638; RTLibExtOPX7TestApplication.cpp:262-280), read where the
 * reference calls GetTexCrd, SampleDiffuse/SampleSpecular/SampleEmission,
 * MeshLight::GetEmission and GetTriangleVNormal (SimpleKernel.h:70-104,
 * 238-272).
 *   texcoord: 6*nt floats (u, v at the triangle's three vertices)
 *   tex_id:   4*nt texture indices: diffuse, specular, emission on hit,
 *             emission when light-sampled; RTPG_TEX_NONE = the reference's
 *             constant white texture (factor 1).  NULL: no texturing.
 *   vnormal:  9*nt floats (vertex normals; glass shades with their
 *             barycentric interpolation, normalised), or NULL (face normal)
 *   texels:   n_texels RGBA8 words (R in the low byte), rows bottom-up
 *   tex:      4*n_tex words per texture: first texel, width, height,
 *             address mode (RTPG_TEX_CLAMP / RTPG_TEX_WRAP); at most 65535
 * Sampling is bilinear on normalised coordinates with 8-bit fractional
 * weights and unorm texels, the filtering of the reference's CUDA texture
 * objects.  A new rtpg_scene_upload drops the attributes. */
int rtpg_scene_attributes(rtpg_context ctx, const float* texcoord, const uint32_t* tex_id, const float* vnormal,
                          const uint32_t* texels, uint64_t n_texels, const uint32_t* tex, uint32_t n_tex);
/* info: [0] BVH nodes, [1] BVH4 depth, [2] stack bound, [3] max leaf size,
 *       [4] lights, [5] triangles.  aabb: world box (lo xyz, hi xyz). */
int rtpg_scene_info(rtpg_context ctx, uint32_t info[6], float aabb[6]);

/* --- camera / frame ---------------------------------------------------------
 * Pinhole camera as RayGenData (SimpleKernel.h:191-199) with u, v, w from
 * Camera::GetUVW (Core/Include/RTLib/Core/Camera.h:78-92). */
int rtpg_camera_set(rtpg_context ctx, const float eye[3], const float u[3], const float v[3], const float w[3]);
int rtpg_frame_resize(rtpg_context ctx, uint32_t width, uint32_t height);
int rtpg_accum_clear(rtpg_context ctx);
/* out: 3*W*H floats (sum of samples, row 0 = bottom), spp: samples accumulated. */
int rtpg_accum_download(rtpg_context ctx, float* out, uint32_t* spp);
/* Display frame of the accumulated image: 4*W*H bytes RGBA (row 0 = bottom),
 * x = sum/spp, x/(1+x), gamma v <= 0.00174 ? 32v : v^(1/2.2), min(int(255.99 g),
 * 255) -- the frame the reference's raygen writes (Test/cuda/SimpleTrace.cu:91-101). */
int rtpg_frame_tonemap(rtpg_context ctx, uint8_t* rgba);

/* --- render -----------------------------------------------------------------
 * Replaces the Params H2D copy + TracerData::Launch -> optixLaunch
 * (RTLibExtOPX7TestApplication.cpp:2191-2198, RTLibExtOPX7Test.h:1345-1348)
 * and the raygen/closest-hit programs (Test/cuda/SimpleTrace.cu,
 * SimpleGuide.cu, SimpleGuide2.cu). */
typedef struct rtpg_render_params {
  uint32_t flags;          /* RTPG_FLAG_* */
  uint32_t spp;            /* samples per pixel in this call */
  uint32_t max_depth;      /* path vertices (reference MaxDepth) */
  uint32_t rr_depth;       /* Russian roulette from this depth (with RTPG_FLAG_RR) */
  uint32_t num_candidates; /* RIS candidates (reference Ris.NumCandidates) */
  uint32_t sample_offset;  /* index of the first sample, for the RNG streams */
  uint64_t seed;
  float bg[3];             /* background radiance (reference MissData.BgColor) */
  float fraction;          /* guide-vs-BSDF mixture (reference SdTree.Fraction) */
  rtpg_sdtree sdtree;      /* guide for RTPG_FLAG_USE_TREE; with RTPG_FLAG_USE_MLP (and not FINAL) the
                              learned guide's training vertices also record into it; else NULL */
  rtpg_hashgrid hashgrid;  /* spatial hash for RTPG_FLAG_USE_GRID */
  rtpg_mqtree mqtree;      /* direction quadtrees for RTPG_FLAG_USE_GRID */
  uint32_t max_paths;      /* paths in flight per chunk (0 = default 32M; at most 64M) */
  rtpg_mlp mlp;            /* learned guide for RTPG_FLAG_USE_MLP, else NULL */
} rtpg_render_params;
int rtpg_render(rtpg_context ctx, const rtpg_render_params* params);
/* Debug frames: the reference's debug programs (__raygen__debug,
 * __closesthit__debug, __miss__debug, SimpleTrace.cu:104-213, 631-734) and
 * DebugFrameType (SimpleKernel.h:160-171).  One camera ray per pixel, jittered
 * as sample params->sample_offset of params->seed, closest hit, then per
 * frame_type: 0 face normal, 1 hit distance, 2 diffuse colour, 3 specular
 * colour, 4 emission, 5 Phong shininess, 6 refraction index, 7 grid value
 * (root weight of the hit cell's building Morton tree; needs RTPG_FLAG_USE_GRID,
 * params->hashgrid, params->mqtree, else 0), 8 world position (no reference
 * counterpart).  value: 4 floats per pixel, the raw quantity (xyz) and the hit
 * primitive's original index as uint32 bits (0xFFFFFFFF on a miss); rgba: the
 * reference's display bytes (normal (n+1)/2, distance t/(1+t), shininess and
 * index x/(1+x), then linear_to_gamma x 255.99 clamped).  Either may be NULL.
 * A miss reads normal (1,0,0), distance 1e20, zeros otherwise.  Pixels in
 * accumulation-buffer order. */
int rtpg_render_debug(rtpg_context ctx, const rtpg_render_params* params, uint32_t frame_type, float* value,
                      uint8_t* rgba);
/* Per-sample relative variance of the last render made with RTPG_FLAG_MOMENTS
 * (at least 2 samples per pixel): the mean over pixels of s^2 / (m^2 + 0.01),
 * s^2 the unbiased variance of that render's samples of y = (r + g + b) / 3
 * and m the pixel's y in the accumulated image (every render since the last
 * clear: a steadier normaliser than the render's own few samples, whose mean
 * would cap the bright outliers that dominate the relative error).  Times the render's cost per sample it ranks
 * estimators by efficiency: the learned guide's fallback decision
 * (rtlib_amd.tracer.render_mlp_guided) compares it between plain, SD-tree-
 * and learned-guide renders.  No reference counterpart (the reference judges
 * methods offline with RTLibExtOPX7Comp.cpp). */
int rtpg_render_relvar(rtpg_context ctx, double* relvar);
/* The raw moments of the last RTPG_FLAG_MOMENTS render: sums[2 i] and
 * sums[2 i + 1] the sum and sum of squares of y over pixel i's samples (in
 * accumulation-buffer order), *spp the render's samples per pixel.  Either
 * may be NULL. */
int rtpg_render_moments(rtpg_context ctx, float* sums, uint32_t* spp);
/* Last render's counters: [0] paths, [1] radiance rays, [2] shadow rays. */
int rtpg_render_stats(rtpg_context ctx, uint64_t stats[3]);
/* Profiling of later renders (no reference counterpart; the reference only
 * times whole frames, RTLibExtOPX7TestApplication.cpp:2527-2532).
 * mode bit 0: time each kernel class with HIP events on the render stream;
 * bit 1: count BVH nodes fetched and triangles tested. */
int rtpg_render_profile(rtpg_context ctx, uint32_t mode);
/* ms[5]: extend, shade, shadow, accumulate, commit of the last render;
 * work[4]: extend nodes, extend triangles, shadow nodes, shadow triangles;
 * launches[5]: kernel launches per class in the last render.  Any pointer may
 * be NULL. */
int rtpg_render_kernel_stats(rtpg_context ctx, double ms[5], uint64_t work[4], uint64_t launches[5]);

/* --- path-replay adjoint ------------------------------------------------------
 * No counterpart in the reference (its integrator is forward only).  For the
 * image estimate I that rtpg_render(params) would produce (same spp, seed,
 * sample offset), adds dLoss/dtheta = sum_pixels dl_dimage . dI/dtheta to
 * grad_albedo (3*nt, original triangle order; Lambertian triangles only) and
 * grad_emission (3 per light, light order of rtpg_scene_lights).  Paths are
 * traced twice with the same random streams (primal, then replay), so memory
 * is O(paths in flight), not O(path length).  Requires no guide, no RIS and no
 * Russian roulette; albedo channels equal to 0 receive no gradient. */
int rtpg_adjoint(rtpg_context ctx, const rtpg_render_params* params, const float* dl_dimage, float* grad_albedo,
                 float* grad_emission);
/* Replace the radiance of every light (3 per light, light order). */
int rtpg_scene_set_emission(rtpg_context ctx, const float* le_rgb);
/* Original triangle index of every light (emissive triangles of positive area). */
int rtpg_scene_lights(rtpg_context ctx, uint32_t* light_tri);

/* --- ray queries (the reference's TraceRadiance / TraceOccluded,
 * SimpleKernel.h:347-366) ----------------------------------------------------
 * rays: 8*n floats (ox, oy, oz, tmin, dx, dy, dz, tmax).
 * hits: 4*n floats (t, u, v, primitive id bits; t = inf on a miss). */
int rtpg_trace_closest(rtpg_context ctx, const float* rays, uint32_t n, float* hits);
int rtpg_trace_occluded(rtpg_context ctx, const float* rays, uint32_t n, uint32_t* occluded);
/* rtpg_trace_closest without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  rays (8*n floats) and hits (4*n floats) are DEVICE
 * pointers owned by the caller, in the layouts of rtpg_trace_closest.  The
 * traversal's spill stack and stack-bound flag use a fixed 32 MiB device array
 * of the library (one per GPU, held until the process ends); rays are traced in
 * launches of as many as its stacks hold, so n is not limited, but a scene
 * whose stack bound leaves room for fewer than 256 rays is refused
 * (RTPG_ERR_LIMIT).  Calls on one GPU share that array: calls that overlap on
 * one GPU (other threads, other contexts' streams) race.  Stream-ordered on the
 * context's stream; returns after the stream has drained. */
int rtpg_trace_closest_noalloc(rtpg_context ctx, const float* rays, uint32_t n, float* hits);

/* --- SD-tree ----------------------------------------------------------------
 * Replaces RTSTreeWrapperT (Utils/OPX7UtilsPathGuiding.h:1333-1767). */
int rtpg_sdtree_create(rtpg_context ctx, const float aabb_min[3], const float aabb_max[3], uint32_t max_dtree_depth,
                       float fraction, rtpg_sdtree* out);
int rtpg_sdtree_destroy(rtpg_sdtree tree);
/* RTSTreeWrapperT::Clear (:1634-1636): one leaf, empty DTrees. */
int rtpg_sdtree_clear(rtpg_sdtree tree);
/* Spatial split threshold: mode 0 = reference size_t(sqrt(2^iter*spp/4)*c)
 * (:1645-1656, c = 4000); mode 1 = c*sqrt(2^iter). */
int rtpg_sdtree_set_split(rtpg_sdtree tree, int mode, float c, int max_mb, float subdiv_th);
/* RTSTreeWrapperT::Reset(iter, spp) (:1645-1656): refine the spatial tree from
 * the recorded weights, then rebuild every building DTree from its sampling tree. */
int rtpg_sdtree_reset(rtpg_sdtree tree, int iter, int samples_per_pass);
/* Octree spatial refinement, the reference's STree2 variant (RTSTreeWrapper2T /
 * RTSTree2T::Refine, SubDivide, :2062-2077, 2108-2142; device side STree2T,
 * :662-783): a leaf whose weight exceeds the threshold splits into 8 children
 * of half the edge, each with 1/8 of the weight (instead of 2 with 1/2).  The
 * octree is stored as three binary levels per split, so lookup, sampling and
 * the recording filters are those of the binary tree.  For the reference's
 * RTSTreeWrapper2T::Reset also set rtpg_sdtree_set_split(tree, 0, 4000, 4000,
 * 0.01).  Takes effect at the next rtpg_sdtree_reset. */
int rtpg_sdtree_set_octree(rtpg_sdtree tree, int enable);
/* RTSTreeWrapperT::Build (:1657-1735): sums, areas, sampling := building. */
int rtpg_sdtree_build(rtpg_sdtree tree);
/* Device record of n samples (DTreeT::RecordIrradiance, :419-427) at
 * positions pos[3n], canonical directions canon[2n]. */
int rtpg_sdtree_record(rtpg_sdtree tree, const float* pos, const float* canon, const float* irradiance,
                       const float* stat_weight, uint32_t n);
/* Recording filters (reference SpatialFilter / DirectionalFilter, :31-43;
 * TraceVertexT::Commit<sFilter, dFilter>, :806-813, 830-866): spatial and
 * directional are 0 (Nearest) or 1 (Box).  Box filters are applied in their
 * stochastic form: each record moves to a uniform point of the reference's box
 * (the voxel of its STree leaf centred on the vertex; the square of edge
 * 0.5^depth centred on its direction) and is recorded there, so every node's
 * expected statistical weight and flux equal the reference's box-filtered
 * values.  The draws are a function of (params.seed, pixel, sample, vertex
 * depth) in rtpg_render, and of (seed, record index, number of earlier
 * rtpg_sdtree_record calls since this one) in rtpg_sdtree_record.  Applies to
 * later recording until changed. */
int rtpg_sdtree_set_filter(rtpg_sdtree tree, int spatial, int directional, uint64_t seed);
/* STree lookup (STreeT::GetDTreeWrapper, :639-648): DTree index per position. */
int rtpg_sdtree_lookup(rtpg_sdtree tree, const float* pos, uint32_t n, uint32_t* dtree);
/* Draw canonical directions from the sampling trees (DTreeNodeT::Sample,
 * :228-277); stream i of PCG32(seed) per sample. */
int rtpg_sdtree_sample(rtpg_sdtree tree, const uint32_t* dtree, uint32_t n, uint64_t seed, float* canon);
/* Sphere density of canonical directions (DTreeT::Pdf, :484-489). */
int rtpg_sdtree_pdf(rtpg_sdtree tree, const uint32_t* dtree, const float* canon, uint32_t n, float* pdf);
/* Sizes: [0] STree nodes, [1] DTrees, [2] building nodes, [3] sampling nodes. */
int rtpg_sdtree_sizes(rtpg_sdtree tree, uint32_t sizes[4]);
/* Flattened export/import (reference Upload/Download, :1353-1615).  Any
 * pointer may be NULL.  stree[s0]; per DTree (n = s1): b_cnt, s_cnt (u32),
 * scal[6*n] = b_sw, b_sum, b_area, s_sw, s_sum, s_area; pools: b_sums[4*s2],
 * b_child[4*s2] (u16), s_sums[4*s3], s_child[4*s3] (u16).  Node pools are
 * concatenated per DTree in DTree order. */
int rtpg_sdtree_export(rtpg_sdtree tree, uint32_t* stree, uint32_t* b_cnt, uint32_t* s_cnt, float* scal,
                       float* b_sums, uint16_t* b_child, float* s_sums, uint16_t* s_child);
int rtpg_sdtree_import(rtpg_sdtree tree, uint32_t n_stree, const uint32_t* stree, uint32_t n_dtrees,
                       const uint32_t* b_cnt, const uint32_t* s_cnt, const float* scal, const float* b_sums,
                       const uint16_t* b_child, const float* s_sums, const uint16_t* s_child);

/* --- SD-tree pass controller ------------------------------------------------
 * Replaces RTSTreeControllerT (Utils/OPX7UtilsPathGuiding.h:1769-1920).
 * state: 0 Record, 1 RecordAndSample, 2 Sample. */
int rtpg_sdtree_ctrl_create(rtpg_sdtree tree, uint32_t sample_for_budget, uint32_t iteration_for_built,
                            float ratio_for_budget, uint32_t sample_per_launch, rtpg_sdtree_ctrl* out);
int rtpg_sdtree_ctrl_destroy(rtpg_sdtree_ctrl c);
int rtpg_sdtree_ctrl_start(rtpg_sdtree_ctrl c);
int rtpg_sdtree_ctrl_beg_trace(rtpg_sdtree_ctrl c);
/* samples: samples traced since BegTrace (0 = sample_per_launch). */
int rtpg_sdtree_ctrl_end_trace(rtpg_sdtree_ctrl c, uint32_t samples);
/* q: [0] state, [1] iteration, [2] samples in the current pass so far,
 *    [3] samples of the current pass, [4] executing (0/1), [5] samples done. */
int rtpg_sdtree_ctrl_query(rtpg_sdtree_ctrl c, uint32_t q[6]);

/* --- hash grid ---------------------------------------------------------------
 * Replaces DoubleBufferedHashGrid3T / DoubleBufferedHashGrid3Buffer
 * (Utils/OPX7UtilsGrid.h:157-353, Test/Include/RTLibExtOPX7Test.h:1472-1570).
 * key_mode 0: the reference's pcg1d-chained bucket and linear-index checksum;
 * key_mode 1: bucket = pcg1d(key) % (slots/32), checksum = key + 1 with key the
 * 30-bit Morton code of the 10-bit cell coordinates (bounds <= 1024 per axis). */
int rtpg_hashgrid_create(rtpg_context ctx, const float aabb_min[3], const float aabb_max[3], const uint32_t bounds[3],
                         uint32_t slots, uint32_t key_mode, rtpg_hashgrid* out);
int rtpg_hashgrid_destroy(rtpg_hashgrid g);
int rtpg_hashgrid_clear(rtpg_hashgrid g);
/* DoubleBufferedHashGrid3Buffer::Update (:1544-1558): previous := current. */
int rtpg_hashgrid_update(rtpg_hashgrid g);
/* insert = 0: FindFromPrv (read-only); insert = 1: FindFromCur (atomicCAS insert).
 * slot = 0xFFFFFFFF when not found / the 32-slot block is full. */
int rtpg_hashgrid_find(rtpg_hashgrid g, const float* pos, uint32_t n, int insert, uint32_t* slot);
/* which = 0: current table, 1: previous table. */
int rtpg_hashgrid_download(rtpg_hashgrid g, int which, uint32_t* checksums);

/* --- Morton direction quadtrees ------------------------------------------
 * Replaces RTMortonQuadTreeWrapperT (Utils/OPX7UtilsMorton.h:590-764) and the
 * device MortonQuadTreeT (:253-449).  Per slot a full pyramid of
 * (4^(level+1)-1)/3 floats; which = 0 building, 1 sampling. */
int rtpg_mqtree_create(rtpg_context ctx, uint32_t slots, uint32_t level, float fraction, rtpg_mqtree* out);
int rtpg_mqtree_destroy(rtpg_mqtree m);
int rtpg_mqtree_allocate(rtpg_mqtree m);  /* zero both pyramids (:608-639) */
int rtpg_mqtree_update(rtpg_mqtree m);    /* swap, zero the new building pyramid (:640-651) */
int rtpg_mqtree_build(rtpg_mqtree m);     /* mortonBuildKernel on the sampling pyramid (OPX7UtilsMortonBuild.cu:2-16) */
int rtpg_mqtree_record(rtpg_mqtree m, const uint32_t* slot, const float* canon, const float* value, uint32_t n);
/* out: 3*n floats (u, v, sphere pdf) */
int rtpg_mqtree_sample(rtpg_mqtree m, const uint32_t* slot, uint32_t n, uint64_t seed, float* out);
int rtpg_mqtree_pdf(rtpg_mqtree m, const uint32_t* slot, const float* canon, uint32_t n, float* pdf);
int rtpg_mqtree_download(rtpg_mqtree m, int which, float* weights);
int rtpg_mqtree_upload(rtpg_mqtree m, int which, const float* weights);
/* info: [0] slots, [1] level, [2] floats per slot */
int rtpg_mqtree_info(rtpg_mqtree m, uint32_t info[3]);

/* --- hash-grid pass controller ---------------------------------------------
 * Replaces RTMortonQuadTreeControllerT (Utils/OPX7UtilsMorton.h:766-993); at
 * the end of each pass it also snapshots the hash grid (previous := current),
 * as the application does (RTLibExtOPX7TestApplication.cpp:2208-2222).
 * state: 0 Record, 1 RecordAndSample, 2 Sample, 3 Locate. */
int rtpg_mqtree_ctrl_create(rtpg_mqtree tree, rtpg_hashgrid grid, uint32_t sample_for_budget,
                            uint32_t iteration_for_built, uint32_t iteration_for_locate, float ratio_for_budget,
                            uint32_t sample_per_launch, rtpg_mqtree_ctrl* out);
int rtpg_mqtree_ctrl_destroy(rtpg_mqtree_ctrl c);
int rtpg_mqtree_ctrl_start(rtpg_mqtree_ctrl c);
int rtpg_mqtree_ctrl_beg_trace(rtpg_mqtree_ctrl c);
int rtpg_mqtree_ctrl_end_trace(rtpg_mqtree_ctrl c, uint32_t samples);
int rtpg_mqtree_ctrl_query(rtpg_mqtree_ctrl c, uint32_t q[6]);

/* --- learned directional guide ------------------------------------------------
 * No counterpart in the reference (its guides are the SD-tree and the Morton
 * hash grid).  A multiresolution hash-grid encoding of the position (16 levels
 * x 2 features, 2^log2_table entries per level, resolutions base_res ..
 * max_res, geometric) concatenated with the outgoing direction, the normal and
 * the roughness feeds a 3 x 64 ReLU MLP (bf16 on the matrix cores, fp32
 * accumulation).  Its 41 outputs give an 8-lobe von Mises-Fisher mixture q and
 * the probability c in [0.05, 0.95] of sampling q instead of the BSDF; the
 * renderer samples the one-sample MIS mixture c q + (1 - c) p_bsdf.  Training
 * minimises mean_i -w_i log(c q(w_i) + (1 - c) p_bsdf(w_i)) with Adam, where
 * w_i = avg_rgb(f Li |cos|) / p_sample of a radiance sample recorded by the
 * tracer (the KL divergence from the radiance-weighted target, up to a
 * constant).  With RTPG_FLAG_USE_MLP a render records training samples from
 * at most record_paths paths of every launch chunk (unless RTPG_FLAG_FINAL):
 * every s-th path, s = ceil(paths / record_paths), starting at path
 * sample_offset mod s, so that they cover the whole image and successive
 * passes record different pixels; with RTPG_FLAG_GUIDE_SAMPLE it samples
 * directions from the guide. */
typedef struct rtpg_mlp_config {
  uint32_t log2_table;   /* hash entries per level (19) */
  float base_res;        /* coarsest grid resolution (16) */
  float max_res;         /* finest grid resolution (256) */
  float lr;              /* Adam step size (1e-2) */
  float beta1, beta2;    /* Adam moments (0.9, 0.99) */
  float eps;             /* Adam epsilon (1e-15) */
  uint32_t batch;        /* samples per training step (2^18) */
  uint32_t capacity;     /* recorded samples kept between trainings (2^20) */
  uint32_t record_paths; /* at most this many paths per render chunk record samples (2^19: a
                            1-spp pass of 1280 x 720 records more than 2^18 samples) */
  uint64_t seed;         /* initialisation */
  uint32_t guide_depth;  /* vertices guided and recorded: depth < guide_depth (0 = all) */
  uint32_t cache_res;    /* 0: the network is evaluated at every guided vertex; n > 0: the
                            renderer reads a spatial cache of n^3 cells x 6 normal directions
                            over the box, refilled from the network after training */
} rtpg_mlp_config;
/* cfg = NULL takes the defaults above.  The position box is made cubic. */
int rtpg_mlp_create(rtpg_context ctx, const float aabb_min[3], const float aabb_max[3], const rtpg_mlp_config* cfg,
                    rtpg_mlp* out);
int rtpg_mlp_destroy(rtpg_mlp m);
/* The spatial cache (cfg.cache_res > 0), refilled first if the network changed:
 * entries = cache_res^3 * 6 (cell (x + res (y + res z)) * 6 + face, face 0 +x,
 * 1 -x, 2 +y, 3 -y, 4 +z, 5 -z); per entry mix[48] = 8 x (mu xyz, kappa),
 * pi[8], pi_j kappa_j / (2 pi (1 - exp(-2 kappa_j)))[8] and c[1].  mix and c
 * may be NULL (only *entries is written). */
int rtpg_mlp_cache_download(rtpg_mlp m, float* mix, float* c, uint32_t* entries);
/* Re-initialise weights and table from seed; clear the Adam state and the recorded samples. */
int rtpg_mlp_reset(rtpg_mlp m, uint64_t seed);
/* Network queries at positions pos[3n] with outgoing directions wo[3n], normals
 * nrm[3n] and roughness rough[n].  raw: 41 outputs per query; mixture: per query
 * 8 x (pi, mu xyz, kappa), then c.  Either may be NULL. */
int rtpg_mlp_forward(rtpg_mlp m, const float* pos, const float* wo, const float* nrm, const float* rough, uint32_t n,
                     float* raw, float* mixture);
/* One direction per query from the vMF mixture of rtpg_mlp_forward's mixture
 * output; out: 4 per query (w xyz, mixture density q(w)); PCG32 stream i of seed. */
int rtpg_mlp_sample(rtpg_mlp m, const float* mixture, uint32_t n, uint64_t seed, float* out);
/* Training samples: pos, wo, nrm (3n each), rough[n], omega[3n] (sampled
 * direction), weight[n] (w_i), pdf_bsdf[n] (BSDF density of omega). */
/* Loss and gradient of the mean loss, without an update.  grad_w: the 14336
 * weights (W0 64x48, W1 64x64, W2 64x64, W3 48x64, row-major out x in);
 * grad_table: 16 * 2^log2_table * 2 (may be NULL). */
int rtpg_mlp_gradient(rtpg_mlp m, const float* pos, const float* wo, const float* nrm, const float* rough,
                      const float* omega, const float* weight, const float* pdf_bsdf, uint32_t n, float* loss,
                      float* grad_w, float* grad_table);
/* `steps` Adam steps on the whole sample set; loss[steps] (may be NULL) before each update. */
int rtpg_mlp_train(rtpg_mlp m, const float* pos, const float* wo, const float* nrm, const float* rough,
                   const float* omega, const float* weight, const float* pdf_bsdf, uint32_t n, uint32_t steps,
                   float* loss);
/* `steps` Adam steps on the samples recorded by renders since the last call
 * (batches of cfg.batch, in recording order), then forget them.  used: the
 * number of recorded samples (may be NULL). */
int rtpg_mlp_train_recorded(rtpg_mlp m, uint32_t steps, float* loss, uint32_t* used);
/* fp32 master copies: weights[14336], table[16 * 2^log2_table * 2]; either may be NULL. */
int rtpg_mlp_export(rtpg_mlp m, float* weights, float* table);
int rtpg_mlp_import(rtpg_mlp m, const float* weights, const float* table);
/* info: [0] weights, [1] table floats, [2] recorded samples, [3] Adam steps
 * taken, [4] levels, [5] entries per level, [6] dense-level mask; res[16]:
 * grid resolution per level (either may be NULL). */
int rtpg_mlp_info(rtpg_mlp m, uint64_t info[7], float res[16]);

#ifdef __cplusplus
}
#endif
#endif /* RTPG_H */
