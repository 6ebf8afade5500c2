// Host-visible launch interface of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cpt_internal.h"

namespace cpt {

// Material record: RTMaterial field order (include/raytracer.h:82-100).
struct DevMaterial {
  float diffuse[3];
  float emissive[3];
  int32_t diffuseTexId, normalTexId, emissiveTexId, matType;
  float transparency, subsurface, specular, metallic, roughness, anisotropic;
  float sheen, sheenTint, clearcoat, clearcoatGloss;
};

// Device-resident scene.  Triangles are in BVH leaf order.
struct DevScene {
  const Node4* nodes4;         // 128-byte 4-wide nodes
  const Node4Q* nodes4q;       // the same nodes, 64-byte quantized
  const float4* tris;          // 3 per triangle: p0 (w = original id bits), e0 = p1-p0, e1 = p2-p0
  const uint4* tri_verts;      // v0, v1, v2, material
  const float4* vpos;          // vertex positions, w = texture u
  const float4* vnrm;          // vertex normals, w = texture v
  const DevMaterial* mats;
  // per triangle in the device tree's leaf order, 8 float4: vertex normals with
  // texture v (3), diffuse + transparency, emissive + specular, (metallic,
  // roughness, diffuse / emissive texture id bits), vertex texture u (3),
  // vertex indices + material (bits)
  const float4* surf;
  const uint32_t* light_tris;  // leaf-order emissive triangles
  uint32_t n_light_tris;
  uint32_t n_tris;
  const float4* bnodes;        // reference binary layout: 2 per node (min, max)
  const uint2* boff;           //   (offset-or-triStart, triN)
  const float4* tris_bin;      //   its triangles, in the reference tree's leaf order
  const float4* tex;           // texels (RGBA) of every texture, concatenated
  const uint4* tex_desc;       // per texture: first texel, width, height
  uint32_t n_tex;
  // environment light (no reference counterpart): env_w x env_h RGB radiance
  // texels (w unused) over directions in latitude-longitude order, row 0 at +y;
  // env_w = 0: none
  const float4* env;
  uint32_t env_w, env_h;
  float qmag;                  // bound on |o + q s| over every slab plane of nodes4q (>= 1e-30)
};

// Per-path SoA state of the wavefront (one slot per live eye path).
struct PathBuf {
  float4* ray_o;     // origin xyz, w = max distance (1e20)
  float4* ray_d;     // direction xyz, w = ray epsilon (1e-4)
  uint32_t* pix;     // pixel index
  uint32_t* sid;     // sample record index (adaptive mode; pixel index otherwise)
  uint32_t* depth;   // bounce count (low 8 bits) | frame offset within the batch << 8
  float4* mul;       // path throughput (pathOutMulTerm)
  float4* smp;       // emitted radiance gathered along the path (pathSample)
  float4* acc;       // xyz accumulated sample, w = sample count bits
  uint64_t* rng;     // PCG32 state
};

// Per-sample records of one frame (adaptive mode): accumulated radiance and
// sample count, and the first-hit G-buffer (pathDirectPos / Norm / Diffuse,
// src/cudaRTBDPTStreamAdapBilatAtomic.cu:172-175,650-655).
struct SampleRec {
  float4* acc;       // xyz radiance, w = N bits
  float4* gpos;
  float4* gnrm;
  float4* gdif;
};

// Shadow (connection) rays: segment + the contribution to add when visible.
struct ShadowBuf {
  float4* ray_o;     // origin, w = max distance
  float4* ray_d;     // direction, w = epsilon
  float4* contrib;   // xyz contribution, w = target slot bits
};

// Light-vertex cache entry (LightVertex, src/cudaRTBDPTStream.cu:19-36).
struct LightVert {
  float pos[3], norm[3], irrad[3], irradDir[3], diff[3], emissive[3];
  float specular, metallic, roughness, pad[3];
};
static_assert(sizeof(LightVert) == 96, "LightVert is 96 bytes");

// One camera view of a multi-view batch: its pixels are [offset, offset + W*H)
// of the batch's concatenated pixel space.
struct ViewDesc {
  uint32_t offset, width, height, pad;
  float cam_pos[3], cam_dir[3], cam_up[3], cam_right[3];
  float fov, tan_half, pad2[2];   // tan_half = tan(fov / 2), computed once on the host
};

struct FrameArgs {
  uint32_t width, height;
  float cam_pos[3], cam_dir[3], cam_up[3], cam_right[3];
  float fov;
  float tan_half;        // tan(fov / 2), computed once on the host
  uint32_t eye_depth, light_depth, light_paths;
  uint32_t flags;        // bit 0: literal shadow endpoint dist - eps
  uint64_t seed;         // stream seed: per-path keys are derived on the device
  uint32_t frame0;       // first frame of the batch (frame == frame0 for single-frame work)
  uint32_t nframes;      // frames traced together as one wavefront
  uint32_t lv_stride;    // light vertices per cache slot
  uint32_t n_views;      // multi-view batch: views (0: one view, width x height)
  const ViewDesc* views; // multi-view batch: pixel p belongs to the last view whose offset <= p
  const uint32_t* frame_dev;  // when set, the batch's first frame is read here (replayed graphs)
};

// Light-vertex cache slots to (re)build: slot s holds the cache of the
// three-frame group starting at frame[i] (cudaRTBDPTStream.cu:875).
constexpr int kLightSlots = 32;
struct LightJobs {
  uint32_t n;
  uint32_t slot[kLightSlots];
  uint32_t frame[kLightSlots];
  // > 0: ignore the lists and (re)build the caches of every group of frames
  // [*fa.frame_dev, + fa.nframes) with up to dev_groups workgroups
  uint32_t dev_groups;
};

// Device counters of one frame (zeroed by frame_begin).
struct FrameCounters {
  uint32_t active[16];            // live paths entering bounce b
  uint32_t shadow[16];            // shadow rays emitted at bounce b
  uint32_t light_n[kLightSlots];  // light-vertex cache size per slot
  uint32_t pad[16];
};

// Cumulative device statistics.
struct DevStats {
  unsigned long long ext_rays, shadow_rays, eye_paths, frames;
  unsigned long long ext_nodes, ext_tris;
  unsigned long long sh_nodes, sh_tris;
  unsigned long long ext_spill, sh_spill;   // stack pushes below the LDS top (k_trace_dyn)
  unsigned long long tail_rays, tail_nodes, tail_tris;   // k_tail: extension rays, nodes and triangles of both kinds
};

constexpr int kMaxDepth = 15;
constexpr int kMaxFilterRadius = 31;   // GAUSSIANCOST_N = 64 entries (include/cudaBilateralFilter.h:3)
constexpr int kBlock = 256;

// ---- launches -------------------------------------------------------------
hipError_t launch_trace(const DevScene& sc, const float4* rays, uint32_t n, const uint32_t* count, uint32_t cull,
                        float4* hits_out, int out_orig, int any_hit, uint32_t layout, DevStats* stats, hipStream_t st, int cost = 0);
hipError_t launch_frame_begin(FrameCounters* fc, uint32_t n_paths, hipStream_t st);
hipError_t launch_light_pass(const DevScene& sc, const FrameArgs& fa, const LightJobs& jobs, LightVert* lv,
                             FrameCounters* fc, hipStream_t st);
// nframes x W*H eye paths, slot = frame offset * W*H + pixel.
// The BVH debug renderer's TraceDepth (spec = 0xFFFFFFFF) or
// TraceDepthParent(spec) per ray over the reference tree: out = 3 u32 per ray.
hipError_t launch_bvh_depth(const DevScene& sc, const float4* rays, uint32_t n, uint32_t spec, uint32_t* out,
                            hipStream_t st);
hipError_t launch_gen_eye(const FrameArgs& fa, PathBuf pb, hipStream_t st);
// Running-mean update of every pixel from the batch's per-frame records, frames in order.
hipError_t launch_apply_frames(const FrameArgs& fa, const float4* rec_acc, float* result, uint32_t* sample_n,
                               int pt, hipStream_t st);
// lds_stack: traversal variant, the stack entries kept in LDS (the rest in
// private memory) and the minimum resident waves per SIMD the register
// allocator is held to: 8, 16, 24 (no minimum), 806, 807, 808 (8 entries, >= 6,
// 7 or 8 waves), 1606 (16 entries, >= 6 waves).  Other values: invalid argument.
inline bool traversal_variant_ok(int v) {
  return v == 8 || v == 16 || v == 24 || v == 806 || v == 807 || v == 808 || v == 1606;
}
hipError_t launch_extend(const DevScene& sc, PathBuf pb, const FrameCounters* fc, int bounce, uint32_t max_n,
                         float4* hits, DevStats* stats, int lds_stack, int quant, hipStream_t st);
// apply: 0 = plain running mean at termination, 1 = running mean + G-buffer
// (adaptive warm-up), 2 = write the sample record only (adaptive frame).
struct Accum {
  float* result;       // W*H*3
  uint32_t* sample_n;  // W*H
  float* gpos;         // W*H*3 running means (adaptive)
  float* gnrm;
  float* gdif;
  float* sq;           // W*H*3 running mean of squared samples (adaptive), or null
  SampleRec rec;
  int apply;
  int variant;         // kVariant* bits
  // path regeneration, first wavefront: per sample slot, the bounce its path
  // ended at and its RNG state then (null: not recorded); per frame of the
  // batch, the last such bounce
  uint32_t* term_bounce;
  uint64_t* term_rng;
  uint32_t* tend;
};
constexpr int kVariantReflOffset = 1;  // reflection offset of AdapBilatAtomic's eye bounce
constexpr int kVariantGbuf = 2;        // first-hit G-buffer
constexpr int kVariantPT = 4;          // path tracing: no connection, total internal reflection reflects
constexpr int kVariantPTRegen = 8;     // the regenerating path tracer's diffuse-to-specular cut
constexpr int kVariantDisney = 16;     // Disney-BRDF path tracer with next-event estimation and MIS
// eye != 0: the bounce's paths are the eye paths k_gen_eye would have written
// to cur (bounce 0 of a fresh batch); they are recomputed from their slot and
// cur is not read.
hipError_t launch_shade(const DevScene& sc, const FrameArgs& fa, PathBuf cur, PathBuf nxt, const float4* hits,
                        FrameCounters* fc, int bounce, uint32_t max_n, ShadowBuf sb, const LightVert* lv,
                        const Accum& ac, hipStream_t st, int eye = 0);
// Whole batch of the plain stream renderer in one kernel: eye paths, every
// bounce's closest hit, shading and connection occlusion per thread; results
// into ac.rec.acc (ac.apply must be 2), counters into fc and stats.
// The reference's "Path Regen" megakernel BDPT (cudaRTBDPTRegen): the
// 640-vertex light cache of frame fa.frame0, and one frame of one thread per
// pixel (eye_depth <= 8 levels) into the running mean.  lv holds
// kRegenLightVerts vertices.
constexpr uint32_t kRegenLightVerts = 640;
hipError_t launch_regen_light(const DevScene& sc, const FrameArgs& fa, LightVert* lv, hipStream_t st);
hipError_t launch_regen_eye(const DevScene& sc, const FrameArgs& fa, const LightVert* lv, float* result,
                            uint32_t* sample_n, int lds_stack, int quant, DevStats* stats, hipStream_t st);
// The bounces from bounce0 on, one thread per live slot of cur, as k_mega
// carries them (grid: workgroups of the persistent launch).
hipError_t launch_tail(const DevScene& sc, const FrameArgs& fa, PathBuf cur, FrameCounters* fc, uint32_t bounce0,
                       uint32_t grid, const LightVert* lv, const Accum& ac, int lds_stack, int quant,
                       DevStats* stats, hipStream_t st);
hipError_t launch_mega(const DevScene& sc, const FrameArgs& fa, FrameCounters* fc, const LightVert* lv,
                       const Accum& ac, int lds_stack, int quant, DevStats* stats, hipStream_t st);
// Path regeneration's second wavefront: T per frame from ac.term_bounce, the
// counters of fc reset, and the restarted paths appended to pb.
hipError_t launch_regen(const FrameArgs& fa, PathBuf pb, const Accum& ac, FrameCounters* fc, hipStream_t st);
hipError_t launch_frame_begin_dev(FrameCounters* fc, const uint32_t* total, uint32_t cap, hipStream_t st);
hipError_t launch_gen_eye_rec(const FrameArgs& fa, PathBuf pb, SampleRec rec, hipStream_t st);

// ---- adaptive sampling (cudaRTBDPTStreamAdapBilatAtomic.cu:1407-1538) ----
struct FilterArgs {
  uint32_t width, height, radius, nd;
  float kc, kp, kn, kd;   // 2 sigma^2 per term
  int use_diff;
  // depth term (no counterpart in the reference filter): kz = 2 sigma_z^2 of the
  // relative depth difference (z_q - z_p) / z_p, z = |first-hit position - cam|;
  // kz = 0 leaves it out
  float kz;
  float cam[3];
};
hipError_t launch_filter_error(const FilterArgs& fa, const float* color, const float* gpos, const float* gnrm,
                               const float* gdif, const float* gauss, const uint32_t* sample_n, float* filtered,
                               float* err, hipStream_t st);
hipError_t launch_error_sum(const float* err, uint32_t n, double* block_sums, float* sum_out, hipStream_t st);
// paths / maxp: per-pixel path counts and the cap a pass's count stops at (maxp 0: no cap)
hipError_t launch_alloc_scan(const float* err, const float* sum, uint32_t n, uint32_t gen, uint32_t* counts,
                             uint32_t* offsets, uint32_t* block_tot, uint32_t* total, uint32_t cap,
                             const uint32_t* paths, uint32_t maxp, hipStream_t st);
// Per-pixel path cap (maxp 0: no-ops): error of capped pixels to 0; counts
// stopped at maxp - paths; paths += counts (or += 1 with counts null).
hipError_t launch_cap_error(float* err, const uint32_t* paths, uint32_t maxp, uint32_t n, hipStream_t st);
hipError_t launch_cap_counts(uint32_t* counts, const uint32_t* paths, uint32_t maxp, uint32_t n, hipStream_t st);
hipError_t launch_add_paths(uint32_t* paths, const uint32_t* counts, uint32_t n, int reset, hipStream_t st);
// Multi-view adaptive batches: view_totals (one per view) become their
// exclusive scan and *total their sum; each pixel's offset then adds its
// view's base.
hipError_t launch_view_bases(uint32_t* view_totals, uint32_t n_views, uint32_t* total, hipStream_t st);
hipError_t launch_offset_views(uint32_t* offsets, uint32_t n, const FrameArgs& fa, const uint32_t* base,
                               hipStream_t st);
hipError_t launch_gen_adaptive(const FrameArgs& fa, const uint32_t* offsets, const uint32_t* counts,
                               const uint32_t* total, uint32_t cap, PathBuf pb, SampleRec rec, hipStream_t st);
hipError_t launch_accum_adaptive(const uint32_t* offsets, const uint32_t* counts, uint32_t n, uint32_t cap,
                                 SampleRec rec, float* result, uint32_t* sample_n, float* gpos, float* gnrm,
                                 float* gdif, float* sq, uint32_t* big_list, uint32_t* big_n, hipStream_t st);

// Empirical squared error (pt_calculateEmpiricalSquareError_kernel, AdapEmp.cu:949-960)
// and the maximum of an error map (thrust::maximum reduce from 0, AdapEmp.cu:1211).
hipError_t launch_emp_error(const float* result, const float* sq, const uint32_t* sample_n, uint32_t n, float* err,
                            hipStream_t st);
hipError_t launch_error_max(const float* err, uint32_t n, float* max_out, hipStream_t st);

// Round-based allocation (pt_genTempAdapPathQueue_kernel rounds + remove_if +
// pt_fillTempAdapPathQueue_kernel; AdapBilatBinary.cu:1348-1388, AdapEmp.cu:1232-1260).
// Round r keeps pixel p unless u * window > max(e_p, min_prob); the first
// kept pixels in pixel order take the remaining queue slots.  Binary: window
// bisected towards gen / nd kept pixels a round; empirical: window = max e.
struct RoundState {
  float left, right, window, maxe;
  uint32_t accum;      // queue slots filled so far
  uint32_t gen;        // slots to fill
  uint32_t done;
  uint32_t rounds;
  uint32_t bisect;     // 1: AdapBilatBinary window bisection, 0: fixed window = max e
  float desired;       // kept pixels per round the bisection aims at
  float min_prob;
  uint32_t pad;
};
hipError_t launch_round_init(RoundState* rs, const float* maxe, uint32_t gen, uint32_t bisect, float desired,
                             float min_prob, uint32_t* counts, uint32_t n, hipStream_t st);
// One round: keep flags, their exclusive scan, count update, state update.
hipError_t launch_round(const FrameArgs& fa, const float* err, uint32_t n, RoundState* rs, uint32_t* flags,
                        uint32_t* rank, uint32_t* block_tot, uint32_t* kept, uint32_t* counts, hipStream_t st);
// Unfilled slots go to pixels 0, 1, ... in order (wrapping past W*H).
hipError_t launch_round_fill(const RoundState* rs, uint32_t n, uint32_t* counts, hipStream_t st);
// Exclusive scan of given per-pixel counts into offsets and their total.
hipError_t launch_scan_counts(const uint32_t* counts, uint32_t n, uint32_t* offsets, uint32_t* block_tot,
                              uint32_t* total, hipStream_t st);
hipError_t launch_shadow(const DevScene& sc, ShadowBuf sb, PathBuf nxt, const FrameCounters* fc, int bounce,
                         uint32_t max_n, int lds_stack, int quant, DevStats* stats, hipStream_t st);
hipError_t launch_frame_end(FrameCounters* fc, DevStats* stats, uint32_t nframes, uint32_t depth, hipStream_t st);
// Closest hit (shadow = 0: hits into out) or connection visibility (shadow = 1:
// contributions into nxt) of the `*count` rays by a persistent grid whose
// lanes refill themselves with the next ray from the global counter `queue`.
// eye (closest hit over quantized nodes only): ray i is the eye ray of path
// slot i of the batch `*eye` describes, computed in place of reading ro / rd.
hipError_t launch_trace_dyn(const DevScene& sc, const float4* ro, const float4* rd, const uint32_t* count,
                            float4* out, const float4* contrib, PathBuf nxt, DevStats* stats, int lds_stack,
                            int quant, int shadow, uint32_t grid, uint32_t refill, uint32_t* queue,
                            hipStream_t st, const FrameArgs* eye = nullptr);
// *frame_dev += k (the device frame counter of replayed graphs).
hipError_t launch_frame_advance(uint32_t* frame_dev, uint32_t k, hipStream_t st);
// Closest hit of the bounce's live paths by a persistent grid: each workgroup
// takes `chunk` rays at a time from the global counter `queue` until the
// queue is drained.
hipError_t launch_extend_persist(const DevScene& sc, PathBuf pb, const FrameCounters* fc, int bounce, float4* hits,
                                 DevStats* stats, int lds_stack, int quant, uint32_t chunk, uint32_t grid,
                                 uint32_t* queue, hipStream_t st);

}  // namespace cpt
