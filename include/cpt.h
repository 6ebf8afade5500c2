/* cpt.h -- C ABI of libcpt_hip.so, the MI355X (gfx950) streaming-BDPT hot path.
 *
 * Plain C types only: opaque handles, plain pointers and sizes.  Every entry
 * point returns 0 on success and a negative CPT_E_* code on failure, never
 * exits, and records a message readable through cpt_last_error() (thread-local).
 * The caller owns every host array; the library owns device memory and one
 * HIP stream per renderer.  Handles are not shared across threads.
 *
 * Reference interfaces each group replaces are cited per declaration, paths
 * relative to the namopth/CudaPT checkout.
 */
#ifndef CPT_H
#define CPT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CPT_OK = 0,
  CPT_E_ARG = -1,      /* bad argument (null handle, size, mode, index) */
  CPT_E_STATE = -2,    /* call out of order (no mesh, no BVH, not uploaded) */
  CPT_E_HIP = -3,      /* a HIP runtime call failed */
  CPT_E_NOMEM = -4,    /* device or host allocation failed */
  CPT_E_LIMIT = -5     /* a size exceeds a documented maximum */
};

typedef struct cpt_scene cpt_scene;
typedef struct cpt_renderer cpt_renderer;

/* RTMaterial, field for field (include/raytracer.h:82-100), 80 bytes. */
typedef struct cpt_material {
  float diffuse[3];
  float emissive[3];
  int32_t diffuse_tex_id, normal_tex_id, emissive_tex_id, mat_type;
  float transparency, subsurface, specular, metallic, roughness, anisotropic;
  float sheen, sheen_tint, clearcoat, clearcoat_gloss;
} cpt_material;

/* One TracePrimitive query (src/cudaRTCommon.cu:66; include/cudaRTCommon.h:61):
 * CURay origin/direction plus the maxDist and rayEpsilon arguments. 32 bytes. */
typedef struct cpt_ray {
  float ox, oy, oz, max_dist;
  float dx, dy, dz, eps;
} cpt_ray;

/* TracePrimitiveResult (include/cudaRTCommon.h:52-59) with the triangle given
 * by its ORIGINAL index (the reference reports the BVH-reordered index).
 * Miss: tri = 0xFFFFFFFF, t = 1e20, u = v = 0.  w = (1 - u) - v. 16 bytes. */
typedef struct cpt_hit {
  uint32_t tri;
  float t, u, v;
} cpt_hit;

/* Last error message of the calling thread ("" when none). */
const char* cpt_last_error(void);

/* Number of visible HIP devices (0 when the runtime finds none). */
int cpt_device_count(int* n);

/* Device memory through the library's own HIP runtime, for callers that hold
 * rays and hits in HBM without a HIP runtime of their own (no reference
 * counterpart: the reference allocates only inside Render). */
int cpt_device_alloc(void** ptr, uint64_t bytes);
int cpt_device_free(void* ptr);
int cpt_memcpy_htod(void* dst, const void* src, uint64_t bytes);
int cpt_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
int cpt_memcpy_dtod(void* dst, const void* src, uint64_t bytes);
int cpt_device_synchronize(void);

/* ---- scene (RTScene, include/raytracer.h:228-290) ---------------------- */

/* RTScene::RTScene (include/raytracer.h:247). */
int cpt_scene_create(cpt_scene** out);

/* Mesh arrays replacing RTScene::m_pVertices / m_pTriangles as filled by
 * RTScene::AddModel (src/raytracer.cpp:238-266).  nrm3 may be NULL (zero
 * normals); uv2 may be NULL.  Invalidates any BVH and device copy. */
int cpt_scene_set_mesh(cpt_scene* s, const float* pos3, const float* nrm3, const float* uv2, uint32_t nv,
                       const uint32_t* idx3, const uint32_t* mat, uint32_t nt);

/* RTScene::m_pMaterials (include/raytracer.h:275). */
int cpt_scene_set_materials(cpt_scene* s, const cpt_material* m, uint32_t n);

/* updateAllSceneMaterialsCudaMem (include/cudaRTCommon.h:95,
 * src/cudaRTCommon.cu:580-599) + updateLightTriCudaMem
 * (src/cudaRTBDPTStream.cu:57-77): pushes the current materials to the device
 * and restarts accumulation of every renderer drawing this scene, as the
 * reference's material dirty flag does (src/cudaRTBDPTStream.cu:845-850). */
int cpt_scene_update_materials(cpt_scene* s);

/* RTScene::m_pTextures (include/raytracer.h:276; loaded by LoadTexture,
 * src/raytracer.cpp:161-179, as float RGBA): appends a width x height texture
 * of 4 floats per texel, row-major, and returns its id, the index materials'
 * diffuseTexId / emissiveTexId refer to.  Sampled as the reference's texture
 * objects are (src/cudaRTCommon.cu:416-441): normalized coordinates, wrap
 * addressing, linear filtering.  Ids outside the added textures mean none.
 * The next render uploads the textures and restarts accumulation. */
int cpt_scene_add_texture(cpt_scene* s, uint32_t width, uint32_t height, const float* rgba, uint32_t* id);
int cpt_scene_clear_textures(cpt_scene* s);

/* No reference counterpart (the reference lights scenes with emissive
 * triangles only, src/cudaRTBDPTStream.cu:57-77): an environment light, a
 * width x height latitude-longitude map of RGB radiance (3 floats per texel,
 * row-major, row 0 looking up +y; u = atan2(d.z, d.x) / 2pi + 1/2 across,
 * v = acos(d.y) / pi down, bilinear, u wrapping).  An eye path that leaves the
 * scene adds the map's radiance along its direction times its throughput;
 * light subpaths still start on emissive triangles only.  width = 0 removes
 * it.  The next render uploads it and restarts accumulation. */
int cpt_scene_set_environment(cpt_scene* s, uint32_t width, uint32_t height, const float* rgb);

/* NPBVHHelper::CreateBVH + CompactBVH::InitialCompactBVH
 * (include/bvhhelper.h:43-54, src/bvhhelper.cpp:55-311): binned SAH, 16 bins,
 * depth <= max_depth (reference default 32), then the triangle reorder of
 * src/raytracer.cpp:267-279 and the 4-wide device re-layout. */
int cpt_bvh_build_sah(cpt_scene* s, uint32_t max_depth);

/* The CompactBVH arrays: bounds6 = node_n x {min xyz, max xyz}, off2 = node_n x
 * {offset-or-triStart, triN}, reorder = nt leaf-order -> original triangle ids.
 * Pass NULL arrays to query node_n only. */
int cpt_bvh_export(cpt_scene* s, float* bounds6, uint32_t* off2, uint32_t* reorder, uint32_t* node_n);

/* No reference counterpart: the most triangles a leaf of the 4-wide device BVH
 * holds (1..64, default 2).  The SAH-optimal collapse of the binary tree picks
 * leaves of at most that many triangles.  Applies to the current BVH (rebuilt at once,
 * uploaded by the next render) and to later builds; hits do not depend on it. */
int cpt_bvh4_set_leaf_size(cpt_scene* s, uint32_t max_tris);

/* Statistics of the 4-wide device BVH: node count, depth, leaf count. */
int cpt_bvh4_info(cpt_scene* s, uint32_t* node_n, uint32_t* depth, uint32_t* leaf_n);

/* initAllSceneCudaMem (include/cudaRTCommon.h:94, src/cudaRTCommon.cu:472-578):
 * device copies of nodes, triangle intersection data, vertices, materials. */
int cpt_scene_upload(cpt_scene* s);

/* freeAllBVHCudaMem (include/cudaRTCommon.h:92) + destruction. */
int cpt_scene_destroy(cpt_scene* s);

/* ---- traversal (TracePrimitive, src/cudaRTCommon.cu:66-127) ------------ */

/* Batched closest hit over host arrays (copies in and out). cull != 0 enables
 * the back-face test of IntersectTri (src/cudaRTCommon.cu:44). */
int cpt_trace_closest(cpt_scene* s, const cpt_ray* rays, uint32_t n, uint32_t cull, cpt_hit* hits);

/* Same over device arrays already resident in HBM, launched on `stream`
 * (a hipStream_t, NULL = the scene's stream).  layout 0 = 4-wide nodes,
 * 1 = the reference's binary CompactBVH layout, 2 = 4-wide nodes quantized to
 * 64 bytes (the same hits: culling stays conservative). */
int cpt_trace_closest_device(cpt_scene* s, const cpt_ray* d_rays, uint32_t n, uint32_t cull, cpt_hit* d_hits,
                             void* stream, uint32_t layout);

/* Occlusion form used by the light-vertex connection
 * (src/cudaRTBDPTStream.cu:685-686): d_occluded[i] = 1 when any triangle lies
 * at 0 <= t < max_dist. Device arrays. */
int cpt_trace_occluded_device(cpt_scene* s, const cpt_ray* d_rays, uint32_t n, uint32_t cull, uint32_t* d_occluded,
                              void* stream);

/* Cross-bilateral filter over device images plus the weighted squared
 * residual: cudaBilateralFilter::updateGaussian + bilaterial_posnorm_kernel /
 * bilaterial_posnormemit_kernel (src/cudaBilateralFilter.cu:5-142) and
 * pt_calculateSquareError_kernel (src/cudaRTBDPTStreamAdapBilatAtomic.cu:1143-1156).
 * Images are width*height x 3 floats; d_diff may be NULL (no albedo term).
 * The spatial Gaussian uses sigma_color, as the reference's callers pass it. */
int cpt_filter_cross_bilateral_device(const float* d_color, const float* d_pos, const float* d_norm,
                                      const float* d_diff, uint32_t width, uint32_t height, float sigma_color,
                                      float sigma_pos, float sigma_norm, float sigma_diff, uint32_t radius,
                                      const uint32_t* d_sample_n, uint32_t nd, float* d_filtered, float* d_err,
                                      void* stream);

/* No counterpart in the reference: cpt_filter_cross_bilateral_device for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments,
 * kernels and results; the spatial Gaussian table, which the entry point above
 * stages in a per-call allocation, lives in a file-scope device array of 64
 * floats instead (radius <= 31; a larger radius is CPT_E_LIMIT).  It calls no
 * hipMalloc, hipFree or allocating library routine.  Every call on one GPU
 * shares that GPU's array and rewrites all of it: calls that overlap on one GPU
 * (two host threads, or two streams) race on the table.  Returns once the
 * filter has run on `stream` (NULL = the null stream): it synchronises that
 * stream, so it is not for use under stream capture. */
int cpt_filter_cross_bilateral_device_noalloc(const float* d_color, const float* d_pos, const float* d_norm,
                                              const float* d_diff, uint32_t width, uint32_t height,
                                              float sigma_color, float sigma_pos, float sigma_norm, float sigma_diff,
                                              uint32_t radius, const uint32_t* d_sample_n, uint32_t nd,
                                              float* d_filtered, float* d_err, void* stream);

/* Traversal cost per ray, the counterpart of the reference's BVH debug
 * renderer (TraceCost / TraceDepth, src/cudaRTCommon.cu:129-273;
 * src/cudaRTBVHDebug.cu): cost2[2i] = nodes visited, cost2[2i+1] = triangles
 * tested by ray i's closest-hit traversal in `layout` (as for
 * cpt_trace_closest_device).  Host arrays. */
int cpt_trace_cost(cpt_scene* s, const cpt_ray* rays, uint32_t n, uint32_t cull, uint32_t layout, uint32_t* cost2);

/* The BVH debug renderer's depth queries (src/cudaRTBVHDebug.cu modes
 * "Normal" and "BVHDepth"), as written over the reference's binary tree (the
 * layout cpt_bvh_export returns), one per ray; each ray's max distance bounds
 * the walk, epsilon is unused.  spec_depth = CPT_DEPTH_NONE: TraceDepth
 * (src/cudaRTCommon.cu:184-231), out3[3i..] = {entered a box, depth, reached
 * a leaf}; otherwise TraceDepthParent(spec_depth) (src/cudaRTCommon.cu:129-182),
 * out3[3i..] = {entered a box, depth or 0xFFFFFFFF when no box at spec_depth was
 * entered, parent node of that box}.  Host arrays. */
#define CPT_DEPTH_NONE 0xFFFFFFFFu
int cpt_trace_depth(cpt_scene* s, const cpt_ray* rays, uint32_t n, uint32_t spec_depth, uint32_t* out3);

/* Node visits and triangle tests of every traversal since the last reset. */
int cpt_trace_counters(cpt_scene* s, uint64_t* nodes, uint64_t* tris, uint64_t* rays, int reset);

/* ---- renderer (RTRenderer / <ns>::Render, include/raytracer.h:292-323,
 *      src/raytracer.cpp:19-31,574-625) ---------------------------------- */

enum {
  CPT_MODE_BDPT_STREAM = 0,          /* cudaRTBDPTStream (src/cudaRTBDPTStream.cu) */
  CPT_MODE_BDPT_ADAPTIVE_PLL = 1,    /* cudaRTBDPTStreamAdapBilatAtomic (src/cudaRTBDPTStreamAdapBilatAtomic.cu):
                                        filter residual, counts proportional to the error */
  CPT_MODE_BDPT_ADAPTIVE_BINARY = 2, /* cudaRTBDPTStreamAdapBilatBinary (src/cudaRTBDPTStreamAdapBilatBinary.cu):
                                        filter residual, rounds with a bisected acceptance window */
  CPT_MODE_BDPT_ADAPTIVE_EMP = 3,    /* cudaRTBDPTStreamAdapEmp (src/cudaRTBDPTStreamAdapEmp.cu): empirical
                                        variance, rounds with the window at the maximum error */
  CPT_MODE_PT_STREAM = 4,            /* cudaRTPTStream (src/cudaRTPTStream.cu): unidirectional path tracing
                                        over the same stream, eye depth 10, a per-frame running mean */
  CPT_MODE_PT_STREAM_REGEN = 5,      /* cudaRTPTStreamRegen (src/cudaRTPTStreamRegen.cu): as 4, and a pixel
                                        whose path ends while the frame is still tracing restarts once */
  CPT_MODE_PT_DISNEY_MIS = 6,        /* no reference counterpart (the reference loads the Disney parameters and
                                        never reads them, src/cudaRTCommon.cu:398-405): a path tracer with
                                        Burley's Disney BRDF (isotropic, opaque), next-event estimation to the
                                        emissive triangles and power-heuristic MIS, Russian roulette from the
                                        fourth bounce, eye depth 8; one sample per path */
  CPT_MODE_BDPT_REGEN = 7            /* cudaRTBDPTRegen (src/cudaRTBDPTRegen.cu), the "Path Regen" baseline: a
                                        megakernel, one thread per pixel; a 640-vertex light cache every third
                                        frame; camera paths restarted for the same pixel when Russian roulette
                                        ends one, at most 5 levels (eye depth, <= 8) */
};

/* RTRenderer::Init(w, h) + SetRendererMode(mode) (src/raytracer.cpp:574-608). */
int cpt_renderer_create(cpt_renderer** out, int mode, uint32_t width, uint32_t height);

/* One <ns>::Render call (src/cudaRTBDPTStream.cu:816-951): one eye path per
 * pixel, light-vertex cache rebuilt on every third frame, running mean updated.
 * Accumulation restarts when (pos, dir, up) differ from the previous call.
 * result (optional, width*height*3 floats, row y = 0 at the bottom) receives
 * the running mean, as the reference's D2H copy (:949). */
int cpt_render_iteration(cpt_renderer* r, cpt_scene* s, const float cam_pos[3], const float cam_dir[3],
                         const float cam_up[3], float fov, float* result);

/* n_frames consecutive Render calls with one host synchronisation at the end;
 * identical results to n_frames calls of cpt_render_iteration. */
int cpt_render_frames(cpt_renderer* r, cpt_scene* s, const float cam_pos[3], const float cam_dir[3],
                      const float cam_up[3], float fov, uint32_t n_frames, float* result);

/* <ns>::GetAttribsN / GetAttribute (src/raytracer.cpp:26-27,
 * include/raytracer.h:32-44).  type: 0 uint, 1 int, 2 bool, 3 float, 4 enum. */
int cpt_renderer_attrib_count(cpt_renderer* r, uint32_t* n);
int cpt_renderer_get_attrib(cpt_renderer* r, uint32_t idx, char* name, uint32_t name_cap, int* type, double* value);
int cpt_renderer_set_attrib(cpt_renderer* r, uint32_t idx, double value);

/* Restart accumulation (frame counter to 0), as a camera change does. */
int cpt_renderer_reset(cpt_renderer* r);

enum {
  CPT_BUF_RESULT = 0,     /* float x3 per pixel: running mean */
  CPT_BUF_SAMPLE_N = 1,   /* uint32 per pixel: accumulated sample count */
  CPT_BUF_GBUF_POS = 2,   /* float x3: first-hit position (adaptive mode) */
  CPT_BUF_GBUF_NORM = 3,  /* float x3: first-hit normal */
  CPT_BUF_GBUF_DIFF = 4,  /* float x3: first-hit albedo */
  CPT_BUF_FILTERED = 5,   /* float x3: cross-bilateral output */
  CPT_BUF_ERROR = 6,      /* float x1: per-pixel error estimate */
  CPT_BUF_SAMPLE_COUNTS = 7, /* uint32: samples drawn per pixel in the last adaptive frame */
  CPT_BUF_ALLOC_ERROR = 8,   /* float x1: the error signal the last allocation used (the learned
                                estimate when a model is attached, else CPT_BUF_ERROR) */
  CPT_BUF_SECOND_MOMENT = 9, /* float x3: running mean of squared samples (squareResult, AdapEmp.cu:927-944) */
  CPT_BUF_PIXEL_PATHS = 10,  /* uint32: eye paths traced per pixel since accumulation restarted (adaptive
                                modes; what "Min Pixel Paths" / "Max Pixel Paths" bound) */
  CPT_BUF_BOUNCE_RAYS = 11   /* uint32 x 32, not per pixel: the last wavefront's extension rays entering
                                bounce 0..15, then its shadow rays emitted at bounce 0..15 (the
                                per-bounce queue lengths, what decides whether refilling finished
                                paths could keep more lanes busy) */
};
int cpt_renderer_read_buffer(cpt_renderer* r, int which, void* out);

/* Work and time counters since the last reset: rays traced (extension +
 * shadow), eye paths completed, frames, and HIP-event milliseconds of the
 * closest-hit kernel, all kernels, and the light pass. */
typedef struct cpt_stats {
  uint64_t extension_rays, shadow_rays, eye_paths, frames;
  uint64_t extend_launches, extend_nodes, extend_tris;
  double extend_ms, shadow_ms, shade_ms, light_ms, total_ms;
  double filter_ms, error_ms, gen_ms, accum_ms;   /* adaptive phases (PERFBREAKDOWN, AdapBilatAtomic.cu:1409-1536) */
  double mlp_ms;                                  /* learned error model: features + inference */
  uint64_t shadow_nodes, shadow_tris;             /* 4-wide nodes visited / triangles tested by shadow rays */
  uint64_t extend_spills, shadow_spills;          /* traversal-stack pushes that went past the LDS entries into
                                                     the private (scratch) stack, refilling-wave kernels */
  uint64_t tail_rays, tail_nodes, tail_tris;      /* "Tail Bounce": extension rays traced by the tail kernel, and
                                                     its nodes / triangles of both ray kinds (the extend_ and
                                                     shadow_ counters above leave them out) */
  double tail_ms;                                 /* the tail kernel's time */
} cpt_stats;
int cpt_renderer_stats(cpt_renderer* r, cpt_stats* out, int reset);

/* Enables per-launch HIP-event timing of every kernel (off by default). */
int cpt_renderer_set_profiling(cpt_renderer* r, int on);

/* <ns>::CleanMem (src/raytracer.cpp:25) + destruction. */
int cpt_renderer_destroy(cpt_renderer* r);

/* ---- Multi-view batch -------------------------------------------------------
 * No counterpart in the reference (one RTRenderer renders one view).  Many
 * camera views of uneven resolution are traced as one wavefront: their pixels
 * are concatenated, every iteration traces frames_in_flight frames of every
 * view, and each view keeps its own running mean (plain streaming BDPT,
 * cudaRTBDPTStream::Render per view).  A view's samples are the ones a
 * one-view renderer of that camera draws (eye streams are keyed by the
 * view-local pixel, the light-vertex caches by frame), so its image equals
 * that renderer's bit for bit. */
typedef struct cpt_batch cpt_batch;
typedef struct cpt_view {
  uint32_t width, height;
  float cam_pos[3], cam_dir[3], cam_up[3];
  float fov;
} cpt_view;
typedef struct cpt_batch_config {
  uint32_t frames_in_flight;  /* frames of every view per iteration, 1..90 */
  uint32_t extend_chunk;      /* 0: one thread per ray; else rays a workgroup takes from the global queue */
  uint32_t extend_workgroups; /* persistent grids: workgroups per compute unit (extend_chunk or refill > 0) */
  uint32_t traversal_variant; /* extension-ray traversal variant: LDS stack entries and minimum waves per
                                 SIMD: 806 (8 entries, >= 6 waves), 807, 808, 1606, 8, 16, 24 */
  uint32_t use_graph;         /* 1: capture one iteration as a hipGraph and replay it */
  uint32_t eye_depth, light_depth, light_paths;
  uint64_t seed;
  uint32_t node_format;       /* 0: 128-byte FP32 4-wide nodes, 1: 64-byte quantized nodes */
  uint32_t refill;            /* extension rays: 0 = as extend_chunk says; 1..64 = persistent waves whose
                                 lanes take new rays once this many are idle (overrides extend_chunk) */
} cpt_batch_config;

/* Defaults: 24 frames in flight, refill 64 with 8 workgroups per CU, variant
 * 806 (shadow rays: the renderer's "Shadow LDS Stack" 806 and "Shadow Refill"
 * 16), quantized nodes, graph on, depths 5/5, 128 light paths, seed
 * 0x5EED0001 (the plain renderer's). */
int cpt_batch_create(cpt_batch** out, const cpt_view* views, uint32_t n_views);
int cpt_batch_get_config(cpt_batch* b, cpt_batch_config* cfg);
/* Restarts accumulation. */
int cpt_batch_set_config(cpt_batch* b, const cpt_batch_config* cfg);
/* Renders `frames` frames of every view (full iterations of frames_in_flight,
 * then one shorter iteration for the rest). */
int cpt_batch_render(cpt_batch* b, cpt_scene* s, uint32_t frames);
int cpt_batch_read_result(cpt_batch* b, uint32_t view, float* out);      /* W*H*3 running mean */
int cpt_batch_sample_counts(cpt_batch* b, uint32_t view, uint32_t* out); /* W*H */
/* One view's buffer (CPT_BUF_*, as cpt_renderer_read_buffer reads it for a
 * one-view renderer): the pixels of that view. */
int cpt_batch_read_buffer(cpt_batch* b, uint32_t view, int which, void* out);
int cpt_batch_stats(cpt_batch* b, cpt_stats* out, int reset);
int cpt_batch_reset(cpt_batch* b);
/* Per-view adaptive sampling.  mode CPT_MODE_BDPT_STREAM: plain frames (the
 * default).  CPT_MODE_BDPT_ADAPTIVE_PLL: every view runs the frames of
 * cudaRTBDPTStreamAdapBilatAtomic::Render (AdapBilatAtomic.cu:1338-1538) -
 * warm-up frames, then per pass of pass_frames frames its own cross-bilateral
 * filter and error sum (or, with `model`, the learned error estimate) and a
 * proportional allocation of pass_frames x its pixel count paths over its own
 * pixels; the paths of all views are traced as one wavefront.  Each view's
 * image equals that of a one-view adaptive renderer of its camera with the
 * same attributes, seed and model, bit for bit.  Replaces the batch's
 * renderer: attributes set before are dropped (set them after), the
 * configuration is kept, accumulation restarts.  Adaptive batches are not
 * autotuned (tune the plain batch first). */
int cpt_batch_set_adaptive(cpt_batch* b, int mode, uint32_t pass_frames, struct cpt_mlp* model);
/* Sets an attribute of the batch's renderer by name (e.g. the adaptive
 * mode's "Filter Radius", "Max Pixel Paths", "Learned Signal"); those the
 * configuration owns (depths, seed, frames in flight, schedule) are refused.
 * Accumulation restarts. */
int cpt_batch_set_attrib(cpt_batch* b, const char* name, double value);
/* Autotuner: times `iterations` iterations of each candidate configuration
 * (extend chunk x workgroups per CU x traversal variant, with the current
 * frames in flight) and keeps the fastest.  *best receives it; ms (may be
 * NULL, n_ms entries) the per-iteration milliseconds of candidates in the
 * order cpt_batch_candidate enumerates them.  Accumulation restarts. */
int cpt_batch_autotune(cpt_batch* b, cpt_scene* s, uint32_t iterations, cpt_batch_config* best, double* ms,
                       uint32_t n_ms);
/* Number of autotuner candidates, and candidate i (the current config with
 * its extend_chunk, extend_workgroups and traversal_variant replaced). */
int cpt_batch_candidates(cpt_batch* b, uint32_t* n);
int cpt_batch_candidate(cpt_batch* b, uint32_t i, cpt_batch_config* cfg);
int cpt_batch_destroy(cpt_batch* b);

/* ---- Learned per-pixel error model ----------------------------------------
 * No counterpart in the reference: CudaPT includes and links FANN
 * (src/cudaRTBDPTStreamBilat.cu:11-12, CudaPT.vcxproj:129) but calls no fann_*
 * function.  The model replaces the filter-residual error estimate of the
 * adaptive renderer (src/cudaRTBDPTStreamAdapBilatAtomic.cu:1143-1156) as the
 * signal the per-pixel sample allocation is proportional to.
 *
 * Network 9 -> 64 -> 64 -> 1, GELU (tanh form) after both hidden layers,
 * BF16 matrix operands with FP32 accumulation, FP32 master weights, Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8).  Parameters are one flat FP32 vector of
 * CPT_MLP_PARAMS entries, row-major with the input index first:
 *   W1[9][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3. */
typedef struct cpt_mlp cpt_mlp;
enum { CPT_MLP_INPUTS = 9, CPT_MLP_HIDDEN = 64, CPT_MLP_PARAMS = 4865 };

/* Creates a model with Glorot-uniform weights drawn from PCG32 stream `seed`
 * (W1, W2, w3 in that order, each entry (2u - 1) * sqrt(6 / (fan_in + fan_out))
 * with u = uniform01 of the stream) and zero biases. */
int cpt_mlp_create(cpt_mlp** out, uint64_t seed);
int cpt_mlp_destroy(cpt_mlp* m);
/* Host copies of the parameters (resets the Adam moments and step). */
int cpt_mlp_set_params(cpt_mlp* m, const float* params, uint32_t n);
int cpt_mlp_get_params(cpt_mlp* m, float* params, uint32_t n);
/* Gradient of the last cpt_mlp_train_step (mean squared error). */
int cpt_mlp_get_grad(cpt_mlp* m, float* grad, uint32_t n);
/* One step over a device batch: x is n x 9 row-major FP32, target n FP32.
 * Computes loss = mean (y - target)^2 and its gradient; with apply != 0 takes
 * one Adam step of rate lr.  *loss (host, may be NULL) receives the loss and
 * makes the call synchronous.  stream NULL = the model's own stream. */
int cpt_mlp_train_step(cpt_mlp* m, const float* d_x, const float* d_target, uint32_t n, float lr, int apply,
                       float* loss, void* stream);
/* No reference counterpart: `steps` training steps with Adam at rate lr over
 * a device data set of n_samples samples (x n_samples x 9, target n_samples),
 * step k taking the batch of samples [j b, (j + 1) b), j = (first_batch + k)
 * mod (n_samples / b); one host call, no synchronisation unless *loss (host,
 * may be NULL) asks for the last step's loss. */
int cpt_mlp_train_steps(cpt_mlp* m, const float* d_x, const float* d_target, uint32_t n_samples, uint32_t batch,
                        uint32_t first_batch, uint32_t steps, float lr, float* loss, void* stream);
/* Inference of n samples into d_y (device FP32). */
int cpt_mlp_infer(cpt_mlp* m, const float* d_x, uint32_t n, float* d_y, void* stream);

/* Per-pixel features of a filter-based adaptive renderer's current state, W*H
 * x 9 FP32 into device memory, with L = Rec.709 luminance, r = running mean,
 * s = running mean of squared samples, f = filtered image and e = filter
 * residual (both recomputed from the current state, as the next pass will),
 * s' = s cross-bilateral filtered with the same guides, v = max(0, mean over
 * channels of s - r^2), v' = max(0, mean of s' - f^2), a = first-hit albedo,
 * n = first-hit normal, p = first-hit position, c = camera:
 *   log1p(L(r)), log1p(L(f)), log(e + 1e-12) / 10, log(v + 1e-12) / 10, L(a),
 *   log(v' + 1e-12) / 10, log1p(|p - c|) / 5, log(max(N, 1)) / 5, n.y. */
int cpt_renderer_mlp_features(cpt_renderer* r, float* d_x);
/* Training target of each pixel against a converged image ref (device,
 * W*H x 3 FP32): log(N * |r - ref|^2 / 3 + 1e-12), an estimate of the log
 * per-sample variance; with clamp > 0 both r and ref are clamped to clamp
 * per channel first (the error of the displayed image). */
int cpt_renderer_mlp_target(cpt_renderer* r, const float* d_ref, float clamp, float* d_t);
/* Attaches a model (NULL detaches): each adaptive frame then allocates its
 * samples proportional to e' = nd / (nd + N) * exp(y) / N (attribute "Learned
 * Signal" 0) or e' = exp(y) / N (1), y the model's output on the features
 * above, in place of the filter-residual estimate.  The model must outlive the
 * attachment. */
int cpt_renderer_set_mlp(cpt_renderer* r, cpt_mlp* m);

#ifdef __cplusplus
}
#endif
#endif /* CPT_H */
