// gfx950 kernels of the streaming-BDPT hot path.
//
//  traverse4      closest/any hit over 128-byte 4-wide nodes, ordered
//                 near-first, stack top 16 entries in LDS (column layout, no
//                 bank conflicts), deeper entries in private memory.
//                 Replaces TracePrimitive (src/cudaRTCommon.cu:66-127).
//  traverse_bin   the same query over the reference's binary CompactBVH
//                 layout (src/cudaRTCommon.cu:527-541), kept as the baseline.
//  k_light_pass   light-vertex cache: pt_genLightPathQueue_kernel +
//                 pt_traceLight_kernel loop (src/cudaRTBDPTStream.cu:180-363,
//                 545-617, 875-906) in one workgroup; stream order of the
//                 reference's stable compaction from a workgroup prefix sum.
//  k_gen_eye      pt_genPathQueue_kernel (src/cudaRTBDPTStream.cu:619-640).
//  k_extend       closest hit of every live path (SoA queue).
//  k_shade        pt_traceSample_kernel body + connection set-up of
//                 pt_connectEyeLightPath_kernel + pt_applyPathQueueResult_kernel
//                 for terminated paths (src/cudaRTBDPTStream.cu:365-543,
//                 665-778); survivors and shadow rays compacted with a wave64
//                 ballot + per-workgroup prefix, one atomic per workgroup.
//  k_shadow       occlusion of the connection segment and the N += 4 update.
//
// Hit semantics: the closest hit is the minimum of (t, original triangle id)
// over all triangles with 0 <= t < maxDist, each tested with IntersectTri's
// exact arithmetic (src/cudaRTCommon.cu:42-64).  Box culling is conservative
// (Ize 2013 slab widening + a slack covering IntersectTri's eps origin shift),
// so the answer does not depend on the node layout or visiting order.
#include <hip/hip_runtime.h>

#include "devmath.h"
#include "kernels.h"

namespace cpt {

// Stack entries per thread in total; the top STK live in LDS, the rest in
// private memory.  A 4-wide level pushes at most 3 entries and the depth is
// <= 31 (checked at build), so 96 entries never overflow.
#define CPT_STACK_TOTAL 96
#define CPT_LDS_STACK 16   // default LDS share (k_light_pass, shadow)

// Smallest float >= 1e-9: float |det| < it  <=>  (double)|det| < M_EPSILON.
__device__ constexpr float kDetEps = 0x1.12e0c0p-30f;
__device__ constexpr float kPiF = 3.14159265358979323846f;
constexpr double kPiD = 3.14159265358979323846;

struct TraceHit {
  float t, u, v;
  uint32_t leaf, orig;
};

// splitmix64 finaliser (Steele, Lea, Flood 2014) keyed by (seed, frame, domain,
// sample): the per-path PCG32 initstate that takes the place of
// curand_init(WangHash(frameN) + index, 0, 0) (src/cudaRTBDPTStream.cu:552,633).
// Domains: 1 eye paths, 2 light paths.
__device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint32_t frame, uint32_t domain, uint32_t sample) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(frame + 1) + ((uint64_t)sample << 32) + ((uint64_t)domain << 60);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}


__device__ __forceinline__ float escape_zero(float v) {
  const float e = 1e-9f;
  if (fabsf(v) < e) v = (v > 0.f) ? v + e : v - e;
  return v;
}
__device__ __forceinline__ float cull_slack(float best, float eps) { return eps + best * 1e-4f + 1e-5f; }

// CURay::IntersectTri (src/cudaRTCommon.cu:42-64); o2 = orig + eps * dir.
__device__ __forceinline__ float isect_tri(f3 o2, f3 d, f3 p0, f3 e0, f3 e1, bool cull, float& u, float& v) {
  if (cull && dot(cross(e0, e1), d) > 0.f) return kInfF;
  f3 de2 = cross(d, e1);
  float div = dot(de2, e0);
  if (fabsf(div) < kDetEps) return kInfF;
  div = 1.0f / div;
  f3 t = o2 - p0;
  f3 te1 = cross(t, e0);
  float rT = dot(te1, e1) * div;
  if (rT < 0.f) return kInfF;
  u = dot(de2, t) * div;
  if (u < 0.f || u > 1.f) return kInfF;
  v = dot(te1, d) * div;
  if (v < 0.f || (u + v) > 1.f) return kInfF;
  return rT;
}

__device__ __forceinline__ bool better(float t, uint32_t oid, const TraceHit& h) {
  return t < h.t || (t == h.t && h.leaf != kInvalid && oid < h.orig);
}

__device__ __forceinline__ void test_leaf(const float4* __restrict__ tris, uint32_t first, uint32_t count, f3 o2, f3 d,
                                          bool cull, TraceHit& h, uint32_t& nt) {
  for (uint32_t i = first; i < first + count; i++) {
    float4 a = tris[3 * i], b = tris[3 * i + 1], c = tris[3 * i + 2];
    float u = 0.f, v = 0.f;
    float t = isect_tri(o2, d, ld3(a), ld3(b), ld3(c), cull, u, v);
    uint32_t oid = __float_as_uint(a.w);
    if (t >= 0.f && better(t, oid, h)) { h.t = t; h.u = u; h.v = v; h.leaf = i; h.orig = oid; }
  }
  nt += count;
}

__device__ __forceinline__ void cswap(uint64_t& a, uint64_t& b) {
  uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo; b = hi;
}

// Child boxes and links of a 4-wide node in the 128-byte FP32 layout (the
// 64-byte quantized layout is tested straight from its words, box4_q).
struct NodeView {
  float lx[4], hx[4], ly[4], hy[4], lz[4], hz[4];
  uint32_t ch[4], cn[4];
};
__device__ __forceinline__ void load_node(const DevScene& sc, uint32_t node, NodeView& v) {
  const Node4* N = sc.nodes4 + node;
  const float4 lx = *(const float4*)N->lox, hx = *(const float4*)N->hix;
  const float4 ly = *(const float4*)N->loy, hy = *(const float4*)N->hiy;
  const float4 lz = *(const float4*)N->loz, hz = *(const float4*)N->hiz;
  const uint4 ch = *(const uint4*)N->child, cn = *(const uint4*)N->count;
  v.lx[0] = lx.x; v.lx[1] = lx.y; v.lx[2] = lx.z; v.lx[3] = lx.w;
  v.hx[0] = hx.x; v.hx[1] = hx.y; v.hx[2] = hx.z; v.hx[3] = hx.w;
  v.ly[0] = ly.x; v.ly[1] = ly.y; v.ly[2] = ly.z; v.ly[3] = ly.w;
  v.hy[0] = hy.x; v.hy[1] = hy.y; v.hy[2] = hy.z; v.hy[3] = hy.w;
  v.lz[0] = lz.x; v.lz[1] = lz.y; v.lz[2] = lz.z; v.lz[3] = lz.w;
  v.hz[0] = hz.x; v.hz[1] = hz.y; v.hz[2] = hz.z; v.hz[3] = hz.w;
  v.ch[0] = ch.x; v.ch[1] = ch.y; v.ch[2] = ch.z; v.ch[3] = ch.w;
  v.cn[0] = cn.x; v.cn[1] = cn.y; v.cn[2] = cn.z; v.cn[3] = cn.w;
}

// Instantiates launch macro M(stack entries, min waves per SIMD) for the
// traversal variant v (traversal_variant_ok); other values return an error.
#define CPT_VARIANT_SWITCH(v, M)            \
  switch (v) {                              \
    case 8: M(8, 0); break;                 \
    case 16: M(16, 0); break;               \
    case 24: M(24, 0); break;               \
    case 806: M(8, 6); break;               \
    case 807: M(8, 7); break;               \
    case 808: M(8, 8); break;               \
    case 1606: M(16, 6); break;             \
    default: return hipErrorInvalidValue;   \
  }

// The top of a thread's traversal stack: its column of a [entries][kBlock] u64
// array in LDS, typed as LDS so that the pop (LDS or private, by depth) stays
// two plain accesses instead of one generic one.
typedef __attribute__((address_space(3))) uint64_t LdsU64;
__device__ __forceinline__ LdsU64* lds_column(uint64_t* shared_base) {
  return (LdsU64*)(shared_base + threadIdx.x);
}

// ---- 4-wide traversal ------------------------------------------------------
// The state of one ray's traversal, advanced one node at a time by ray_step, so
// that a lane whose ray has finished can start another while the rest of its
// wave is still traversing (k_trace_dyn) or simply loop (traverse4).
// pn, pf: -o * idir moved out by the quantized box test's slack, per axis.
struct RayState {
  f3 o, d, idir, o2, pn, pf;
  float eps;
  TraceHit h;
  uint32_t node;
  int sp;
};
// qmag bounds the magnitude of every quantized slab plane of the scene.
__device__ __forceinline__ void ray_begin(RayState& R, f3 o, f3 d, float maxd, float eps, float qmag) {
  R.o = o;
  R.d = d;
  R.idir = mk3(1.f / escape_zero(d.x), 1.f / escape_zero(d.y), 1.f / escape_zero(d.z));
  // Slack of the quantized box test (box4_q): 2^-20 = 16 u per unit of
  // |idir| (qmag + |o|), twice the rounding bound of that test's distances.
  const float k = 0x1p-20f;
  const float sx = (fabsf(R.idir.x) * (qmag + fabsf(o.x))) * k;
  const float sy = (fabsf(R.idir.y) * (qmag + fabsf(o.y))) * k;
  const float sz = (fabsf(R.idir.z) * (qmag + fabsf(o.z))) * k;
  const f3 oi = mk3(-(o.x * R.idir.x), -(o.y * R.idir.y), -(o.z * R.idir.z));
  R.pn = mk3(oi.x - sx, oi.y - sy, oi.z - sz);
  R.pf = mk3(oi.x + sx, oi.y + sy, oi.z + sz);
  R.o2 = o + d * eps;
  R.eps = eps;
  R.h.t = maxd; R.h.u = R.h.v = 0.f; R.h.leaf = kInvalid; R.h.orig = kInvalid;
  R.node = 0;
  R.sp = 0;
}
// The four child boxes of a 64-byte quantized node against the ray: entry and
// exit distances tn, tf (entry clamped at 0), children and leaf counts.  A
// slab plane is o_a + q s_a with s_a a power of two, so its distance
// (o_a + q s_a - ray.o_a) idir_a is fma(q, s_a idir_a, c_a) with
// c_a = fma(o_a, idir_a, -ray.o_a idir_a) per node (s_a idir_a is exact); the
// near plane of each axis (by the sign of idir_a) gives the entry and the far
// plane the exit, so no per-axis min / max is needed.  With M = sc.qmag,
// u = 2^-24 and slack = 16 u |idir_a| (M + |ray.o_a|), the distance differs
// from the exact one to the host's plane by at most six roundings (idir
// against 1 / d, the plane's own, -ray.o_a idir_a, the slack's addition and
// the two fmas), each within u (|idir_a| (M + |ray.o_a|) + slack): under half
// the slack, which is subtracted from entries (R.pn) and added to exits
// (R.pf), so no box that the ray meets in exact arithmetic is culled.
__device__ __forceinline__ void box4_q(const DevScene& sc, const RayState& R, float tn[4], float tf[4],
                                       uint32_t ch[4], uint32_t cn[4]) {
  const uint4* N = (const uint4*)(sc.nodes4q + R.node);
  const uint4 A = N[0], B = N[1], C = N[2], D = N[3];
  const float sx = __uint_as_float((A.w & 0xFFu) << 23) * R.idir.x;
  const float sy = __uint_as_float(((A.w >> 8) & 0xFFu) << 23) * R.idir.y;
  const float sz = __uint_as_float(((A.w >> 16) & 0xFFu) << 23) * R.idir.z;
  const float ox = __uint_as_float(A.x), oy = __uint_as_float(A.y), oz = __uint_as_float(A.z);
  const float nx = __builtin_fmaf(ox, R.idir.x, R.pn.x), fx = __builtin_fmaf(ox, R.idir.x, R.pf.x);
  const float ny = __builtin_fmaf(oy, R.idir.y, R.pn.y), fy = __builtin_fmaf(oy, R.idir.y, R.pf.y);
  const float nz = __builtin_fmaf(oz, R.idir.z, R.pn.z), fz = __builtin_fmaf(oz, R.idir.z, R.pf.z);
  const bool px = R.idir.x >= 0.f, py = R.idir.y >= 0.f, pz = R.idir.z >= 0.f;
  const uint32_t qnx = px ? B.x : B.y, qfx = px ? B.y : B.x;
  const uint32_t qny = py ? B.z : B.w, qfy = py ? B.w : B.z;
  const uint32_t qnz = pz ? C.x : C.y, qfz = pz ? C.y : C.x;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int sh = 8 * c;
    const float ex = __builtin_fmaf((float)((qnx >> sh) & 0xFFu), sx, nx);
    const float ey = __builtin_fmaf((float)((qny >> sh) & 0xFFu), sy, ny);
    const float ez = __builtin_fmaf((float)((qnz >> sh) & 0xFFu), sz, nz);
    const float lx = __builtin_fmaf((float)((qfx >> sh) & 0xFFu), sx, fx);
    const float ly = __builtin_fmaf((float)((qfy >> sh) & 0xFFu), sy, fy);
    const float lz = __builtin_fmaf((float)((qfz >> sh) & 0xFFu), sz, fz);
    tn[c] = fmaxf(fmaxf(ex, ey), fmaxf(ez, 0.f));
    tf[c] = fminf(fminf(lx, ly), lz);
  }
  cn[0] = C.z & 0xFFFFu; cn[1] = C.z >> 16; cn[2] = C.w & 0xFFFFu; cn[3] = C.w >> 16;
  ch[0] = D.x; ch[1] = D.y; ch[2] = D.z; ch[3] = D.w;
}

// The four child boxes of node R.node (either layout): entry / exit
// distances, children and leaf counts.
template <bool Q>
__device__ __forceinline__ void child_boxes(const DevScene& sc, const RayState& R, float tn[4], float tf[4],
                                            uint32_t ch[4], uint32_t cn[4]) {
  if constexpr (Q) {
    box4_q(sc, R, tn, tf, ch, cn);
  } else {
    NodeView nv;
    load_node(sc, R.node, nv);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      float t0x = (nv.lx[c] - R.o.x) * R.idir.x, t1x = (nv.hx[c] - R.o.x) * R.idir.x;
      float t0y = (nv.ly[c] - R.o.y) * R.idir.y, t1y = (nv.hy[c] - R.o.y) * R.idir.y;
      float t0z = (nv.lz[c] - R.o.z) * R.idir.z, t1z = (nv.hz[c] - R.o.z) * R.idir.z;
      tn[c] = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.f));
      tf[c] = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * 1.0000004f;
      ch[c] = nv.ch[c];
      cn[c] = nv.cn[c];
    }
  }
}

// The leaf children's triangles as one run (child order, triangle order):
// child c contributes ln[c] triangles from lf[c] on.  Run position k lies in
// child c when p_c <= k < p_(c+1) (prefix sums of the counts); its triangle is
// k + (first_c - p_c): selects, no branches.  Returns true when an occlusion
// ray (ANY) found a hit.
template <bool ANY>
__device__ __forceinline__ bool test_run(const DevScene& sc, RayState& R, bool cull, const uint32_t lf[4],
                                         const uint32_t ln[4], uint32_t& nt) {
  const uint32_t p1 = ln[0], p2 = p1 + ln[1], p3 = p2 + ln[2];
  const uint32_t ntri = p3 + ln[3];
  nt += ntri;
  const uint32_t d0 = lf[0], d1 = lf[1] - p1, d2 = lf[2] - p2, d3 = lf[3] - p3;
  // All of a triangle's 48 bytes are fetched at once (the empty asm needs the
  // vertex before the determinant test, which would otherwise delay its load
  // to a second round trip); one triangle per round trip for both ray kinds:
  // fetching two for occlusion rays spilled five registers of the shadow
  // kernel and ran it 2 % slower end to end.
  for (uint32_t k = 0; k < ntri; k++) {
    const uint32_t i = k + (k < p1 ? d0 : k < p2 ? d1 : k < p3 ? d2 : d3);
    const float4 a = sc.tris[3 * i], b = sc.tris[3 * i + 1], cc = sc.tris[3 * i + 2];
    asm volatile("" : : "v"(a.x), "v"(a.y), "v"(a.z));
    float u = 0.f, v = 0.f;
    const float t = isect_tri(R.o2, R.d, ld3(a), ld3(b), ld3(cc), cull, u, v);
    const uint32_t oid = __float_as_uint(a.w);
    if (t >= 0.f && better(t, oid, R.h)) { R.h.t = t; R.h.u = u; R.h.v = v; R.h.leaf = i; R.h.orig = oid; }
    if (ANY && R.h.leaf != kInvalid) return true;
  }
  return false;
}

// Pops the nearest pending node still in front of the best hit into R.node;
// false when the stack runs out (the ray is finished).
template <int STK>
__device__ __forceinline__ bool pop_node(RayState& R, const LdsU64* lstk, const uint64_t* pstk) {
  while (R.sp > 0) {
    R.sp--;
    uint64_t e = R.sp < STK ? lstk[R.sp * kBlock] : pstk[R.sp - STK];
    float tn = __uint_as_float((uint32_t)(e >> 32));
    if (tn <= R.h.t + cull_slack(R.h.t, R.eps)) { R.node = (uint32_t)e; return true; }
  }
  return false;
}

// One node: the four child boxes against the ray (culled beyond the best hit
// plus slack); the leaf children hit are then tested as one run of triangles
// (test_run), so that lanes whose leaves sit under different children test
// them in the same loop; interior children are culled again with the updated
// best hit, sorted near-first, the nearest taken next and the rest pushed
// (LDS stack top, private stack below it).  The closest hit is the
// (t, original id) minimum, which neither the test order nor the conservative
// culling changes.  Returns true when the ray is finished.
// lstk points at this thread's column of a [STK][kBlock] u64 array in LDS.
// spill (an LDS word of the wave, or null) counts pushes that land below the
// LDS top, in the private (scratch) stack: an LDS atomic on that rare event
// instead of a counter register in the loop.
template <bool ANY, int STK, bool Q>
__device__ __forceinline__ bool ray_step(const DevScene& sc, RayState& R, bool cull, LdsU64* lstk, uint64_t* pstk,
                                         uint32_t& nn, uint32_t& nt, uint32_t* spill) {
  float tn[4], tf[4];
  uint32_t ch[4], cn[4];
  child_boxes<Q>(sc, R, tn, tf, ch, cn);
  nn++;
  const float lim = R.h.t + cull_slack(R.h.t, R.eps);
  uint64_t key[4];
  uint32_t lf[4], ln[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const float tnear = tn[c];
    bool hit = ch[c] != kInvalid && tnear <= tf[c] && tnear <= lim;
    const bool leaf = cn[c] > 0;
    lf[c] = ch[c];
    ln[c] = hit && leaf ? cn[c] : 0u;
    // 0 marks no interior child (a child link is never the root, 0)
    key[c] = hit && !leaf ? ((uint64_t)__float_as_uint(tnear) << 32) | ch[c] : 0ull;
  }
  if (ln[0] + ln[1] + ln[2] + ln[3]) {
    if (test_run<ANY>(sc, R, cull, lf, ln, nt)) return true;
    const float lim2 = R.h.t + cull_slack(R.h.t, R.eps);
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (__uint_as_float((uint32_t)(key[q] >> 32)) > lim2) key[q] = 0ull;
  }
  // Near-first for both ray kinds: occlusion rays taken in node order visit
  // 15 % fewer nodes but test 23 % more triangles and run 1 % slower on the
  // sphere grid.  Sorted ascending, the n interior children hit are the last n
  // keys: key[4 - n] is taken next and key[3], key[2], key[1] (those of them
  // past it) are pushed farthest first.  Slots from sp on are free, so when
  // all three fit the LDS top they are written unconditionally and sp moves
  // by n - 1.
  // With one interior child hit there is nothing to sort or push: its key is
  // the OR of the four (the others are 0).
  const int n = ((uint32_t)key[0] != 0u) + ((uint32_t)key[1] != 0u) + ((uint32_t)key[2] != 0u) +
                ((uint32_t)key[3] != 0u);
  if (n == 1) {
    R.node = (uint32_t)(key[0] | key[1] | key[2] | key[3]);
    return false;
  }
  if (n) {
    cswap(key[0], key[1]); cswap(key[2], key[3]); cswap(key[0], key[2]); cswap(key[1], key[3]); cswap(key[1], key[2]);
    R.node = n >= 3 ? (uint32_t)(n >= 4 ? key[0] : key[1]) : (uint32_t)key[2];
    if (R.sp + 3 <= STK) {
      lstk[R.sp * kBlock] = key[3];
      lstk[(R.sp + 1) * kBlock] = key[2];
      lstk[(R.sp + 2) * kBlock] = key[1];
      R.sp += n - 1;
    } else {
#pragma unroll
      for (int c = 3; c >= 1; c--) {
        if (c + n > 4 && R.sp < CPT_STACK_TOTAL) {
          if (R.sp < STK) {
            lstk[R.sp * kBlock] = key[c];
          } else {
            pstk[R.sp - STK] = key[c];
            if (spill) atomicAdd(spill, 1u);
          }
          R.sp++;
        }
      }
    }
    return false;
  }
  return !pop_node<STK>(R, lstk, pstk);
}

// Closest (ANY: any) hit of one ray, the loop of ray_step.
// Q: nodes in the 64-byte quantized layout.
template <bool ANY, int STK = CPT_LDS_STACK, bool Q = false>
__device__ bool traverse4(const DevScene& sc, f3 o, f3 d, float maxd, float eps, bool cull, LdsU64* lstk,
                          TraceHit& h, uint32_t& nn, uint32_t& nt) {
  RayState R;
  ray_begin(R, o, d, maxd, eps, sc.qmag);
  uint64_t pstk[CPT_STACK_TOTAL - STK];
  while (!ray_step<ANY, STK, Q>(sc, R, cull, lstk, pstk, nn, nt, nullptr)) {
  }
  h = R.h;
  return h.leaf != kInvalid;
}

// Same query over the reference's binary layout (right child popped first).
template <bool ANY>
__device__ bool traverse_bin(const DevScene& sc, f3 o, f3 d, float maxd, float eps, bool cull, TraceHit& h,
                             uint32_t& nn, uint32_t& nt) {
  const f3 idir = mk3(1.f / escape_zero(d.x), 1.f / escape_zero(d.y), 1.f / escape_zero(d.z));
  const f3 o2 = o + d * eps;
  h.t = maxd; h.u = h.v = 0.f; h.leaf = kInvalid; h.orig = kInvalid;
  uint32_t stk[64];
  int sp = 0;
  stk[0] = 0;
  while (sp >= 0) {
    uint32_t n = stk[sp--];
    float4 mn = sc.bnodes[2 * n], mx = sc.bnodes[2 * n + 1];
    uint2 off = sc.boff[n];
    nn++;
    float t0x = (mn.x - o.x) * idir.x, t1x = (mx.x - o.x) * idir.x;
    float t0y = (mn.y - o.y) * idir.y, t1y = (mx.y - o.y) * idir.y;
    float t0z = (mn.z - o.z) * idir.z, t1z = (mx.z - o.z) * idir.z;
    float tnear = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.f));
    float tfar = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * 1.0000004f;
    if (!(tnear <= tfar && tnear <= h.t + cull_slack(h.t, eps))) continue;
    if (off.y == 0) {
      if (mn.x > mx.x) continue;                     // empty leaf of a degenerate split
      if (sp < 62) { stk[++sp] = n + 1; stk[++sp] = n + off.x; }
    } else {
      test_leaf(sc.tris_bin, off.x, off.y, o2, d, cull, h, nt);
      if (ANY && h.leaf != kInvalid) return true;
    }
  }
  return h.leaf != kInvalid;
}

// ---- the BVH debug renderer's depth queries ---------------------------------
// CURay::IntersectAABB (src/cudaRTCommon.cu:24-40) as written: the entry
// distance (0 within M_EPSILON of the origin), kInfF on a miss.
__device__ __forceinline__ float ref_aabb(f3 o, f3 rcp, float4 mn, float4 mx) {
  const float t0x = (mn.x - o.x) * rcp.x, t1x = (mx.x - o.x) * rcp.x;
  const float t0y = (mn.y - o.y) * rcp.y, t1y = (mx.y - o.y) * rcp.y;
  const float t0z = (mn.z - o.z) * rcp.z, t1z = (mx.z - o.z) * rcp.z;
  const float minmax = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  const float maxmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
  if (minmax >= maxmin && (minmax > 0.f || maxmin > 0.f)) return ((double)maxmin > 1e-9) ? maxmin : 0.f;
  return kInfF;
}
// TraceDepth / TraceDepthParent (src/cudaRTCommon.cu:129-231) as written, over
// the reference's binary tree: right child popped first, a 32-entry command
// stack, at most 2048 pops.  spec == 0xFFFFFFFF: TraceDepth, out = {hit,
// depth, leaf}; otherwise TraceDepthParent(spec), out = {hit, depth
// (0xFFFFFFFF for -1), parent node}.
__global__ __launch_bounds__(kBlock) void k_bvh_depth(DevScene sc, const float4* __restrict__ rays, uint32_t n,
                                                      uint32_t spec, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 a = rays[2 * i], b = rays[2 * i + 1];
  const f3 o = ld3(a), d = ld3(b);
  const f3 rcp = mk3(1.f / escape_zero(d.x), 1.f / escape_zero(d.y), 1.f / escape_zero(d.z));
  float minI = a.w;
  uint32_t cmd[32], par[32], dep[32];
  cmd[0] = 0; par[0] = 0; dep[0] = 0;
  int sp = 0;
  uint32_t traced = 0;
  uint32_t r0 = 0, r1 = 0, r2 = 0;
  if (spec == 0xFFFFFFFFu) {
    uint32_t depth = 0;
    bool leaf = false;
    while (sp >= 0 && traced++ < 2048u) {
      const uint32_t nd = cmd[sp--];
      const float tb = ref_aabb(o, rcp, sc.bnodes[2 * nd], sc.bnodes[2 * nd + 1]);
      if (tb >= 0.f && tb < minI) {
        const uint2 off = sc.boff[nd];
        if (off.y == 0) {
          if (sp < 30) { cmd[++sp] = nd + 1; cmd[++sp] = nd + off.x; }
        } else {
          leaf = true; minI = tb; depth = (uint32_t)(sp + 2);
        }
      } else {
        depth = max(depth, (uint32_t)(sp + 1));
      }
    }
    r0 = depth != 0; r1 = depth; r2 = leaf;
  } else {
    int depth = -1;
    uint32_t parent = 0;
    bool any = false;
    while (sp >= 0 && traced++ < 2048u) {
      const int cp = sp--;
      const uint32_t nd = cmd[cp];
      const float tb = ref_aabb(o, rcp, sc.bnodes[2 * nd], sc.bnodes[2 * nd + 1]);
      if (tb >= 0.f && tb < minI) {
        any = true;
        const uint32_t cd = dep[cp];
        if (spec == cd) { minI = tb; depth = (int)cd; parent = par[cp]; break; }
        const uint2 off = sc.boff[nd];
        if (off.y == 0 && sp < 30) {
          ++sp; cmd[sp] = nd + 1; par[sp] = nd; dep[sp] = cd + 1;
          ++sp; cmd[sp] = nd + off.x; par[sp] = nd; dep[sp] = cd + 1;
        }
      }
    }
    r0 = any; r1 = any ? (uint32_t)depth : 0xFFFFFFFFu; r2 = any ? parent : 0u;
  }
  out[3 * i] = r0;
  out[3 * i + 1] = r1;
  out[3 * i + 2] = r2;
}
hipError_t launch_bvh_depth(const DevScene& sc, const float4* rays, uint32_t n, uint32_t spec, uint32_t* out,
                            hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (!rays || !out || !sc.bnodes || !sc.boff) return hipErrorInvalidValue;
  k_bvh_depth<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(sc, rays, n, spec, out);
  return hipGetLastError();
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Host-side guards: a launch never receives a null buffer it will index.
static bool pathbuf_ok(const PathBuf& b) {
  return b.ray_o && b.ray_d && b.pix && b.sid && b.depth && b.mul && b.smp && b.acc && b.rng;
}
static bool rec_ok(const SampleRec& r) { return r.acc && r.gpos && r.gnrm && r.gdif; }

// ---------------------------------------------------------------- trace API
// out mode 0: cpt_hit {orig id, t, u, v}; 1: {t, leaf bits, u, v}; 2: occluded flag.
// MINW: minimum resident waves per SIMD requested from the register allocator
// (0 = compiler default).
template <int LAYOUT, bool ANY, int OUT, int STK = CPT_LDS_STACK, int MINW = 0>
__global__ __launch_bounds__(kBlock, MINW) void k_trace(DevScene sc, const float4* __restrict__ ro,
                                                  const float4* __restrict__ rd, uint32_t stride,
                                                  const uint32_t* __restrict__ countp, uint32_t n, uint32_t cull,
                                                  float4* __restrict__ out, DevStats* __restrict__ stats) {
  __shared__ uint64_t lstk[STK * kBlock];
  uint32_t count = countp ? *countp : n;
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t nn = 0, nt = 0;
  if (i < count) {
    float4 a = ro[(size_t)i * stride], b = rd[(size_t)i * stride];
    TraceHit h;
    bool hit = LAYOUT == 0   ? traverse4<ANY, STK>(sc, ld3(a), ld3(b), a.w, b.w, cull != 0, lds_column(lstk), h, nn, nt)
               : LAYOUT == 2 ? traverse4<ANY, STK, true>(sc, ld3(a), ld3(b), a.w, b.w, cull != 0, lds_column(lstk), h,
                                                         nn, nt)
                             : traverse_bin<ANY>(sc, ld3(a), ld3(b), a.w, b.w, cull != 0, h, nn, nt);
    if (OUT == 0) {
      float4 r;
      r.x = __uint_as_float(hit ? h.orig : kInvalid);
      r.y = hit ? h.t : kInfF;
      r.z = hit ? h.u : 0.f;
      r.w = hit ? h.v : 0.f;
      out[i] = r;
    } else if (OUT == 1) {
      float4 r;
      r.x = hit ? h.t : kInfF;
      r.y = __uint_as_float(hit ? h.leaf : kInvalid);
      r.z = h.u;
      r.w = h.v;
      out[i] = r;
    } else if (OUT == 2) {
      ((uint32_t*)out)[i] = hit ? 1u : 0u;
    } else {
      ((uint2*)out)[i] = make_uint2(nn, nt);      // this ray's node visits and triangle tests
    }
  }
  if (stats) {
    uint32_t sn = wave_sum(nn), st = wave_sum(nt);
    if ((threadIdx.x & 63) == 0 && (sn | st)) {
      atomicAdd(&stats->ext_nodes, (unsigned long long)sn);
      atomicAdd(&stats->ext_tris, (unsigned long long)st);
    }
  }
}

hipError_t launch_trace(const DevScene& sc, const float4* rays, uint32_t n, const uint32_t* count, uint32_t cull,
                        float4* hits_out, int out_orig, int any_hit, uint32_t layout, DevStats* stats,
                        hipStream_t st, int cost) {
  if (n == 0) return hipSuccess;
  dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
  const float4* ro = rays;
  const float4* rd = rays + 1;
  if (layout > 2 || (layout == 2 && !sc.nodes4q) || (layout == 0 && !sc.nodes4)) return hipErrorInvalidValue;
#define CPT_TR(L, A, O) k_trace<L, A, O><<<grid, block, 0, st>>>(sc, ro, rd, 2, count, n, cull, hits_out, stats)
  if (cost) {
    if (layout == 0) CPT_TR(0, false, 3); else if (layout == 2) CPT_TR(2, false, 3); else CPT_TR(1, false, 3);
  } else if (any_hit) {
    if (layout == 0) CPT_TR(0, true, 2); else if (layout == 2) CPT_TR(2, true, 2); else CPT_TR(1, true, 2);
  } else if (out_orig) {
    if (layout == 0) CPT_TR(0, false, 0); else if (layout == 2) CPT_TR(2, false, 0); else CPT_TR(1, false, 0);
  } else {
    if (layout == 0) CPT_TR(0, false, 1); else if (layout == 2) CPT_TR(2, false, 1); else CPT_TR(1, false, 1);
  }
#undef CPT_TR
  return hipGetLastError();
}

// ---------------------------------------------------------------- shading
// src/cudaRTCommon.cu:284-365 with the reference's mixed float/double
// expressions (M_PI and the bias epsilons are double literals there).
__device__ __forceinline__ bool prob_rand(uint64_t& st, uint64_t inc, float prob) {
  if ((double)prob < kBiasEps) return false;
  if ((double)prob + kBiasEps > 1.0) return true;
  return uniform01(st, inc) < prob;
}
__device__ __forceinline__ float vis_smith(float r, float nov, float nol) {
  float k = (r * r) / 2.f;
  return (nov / (nov * (1.f - k) + k)) * (nol / (nol * (1.f - k) + k));
}
__device__ __forceinline__ float d_ggx(float r, float noh) {
  float m = r * r, m2 = m * m;
  float d = (noh * m2 - noh) * noh + 1.f;
  return (float)((double)m2 / (kPiD * (double)d * (double)d));
}
// x^5 of Schlick's term: the power in double (x^2 exact, x^4 and x^5 rounded
// at 2^-53) rounded once to float, the correctly rounded x^5 except within
// ~2^-51 of a rounding boundary (the reference's powf is within an ulp of it);
// a few instructions where a general powf takes dozens.
__device__ __forceinline__ float pow5f(float x) {
  const double d = (double)x, d2 = d * d;
  return (float)(d2 * d2 * d);
}
__device__ __forceinline__ f3 f_schlick(f3 s, float voh) {
  float fc = pow5f(1.f - voh);
  double z = 50.0 * (double)s.z;
  float first = (float)(z < 0.0 ? 0.0 : (z > 1.0 ? 1.0 : z)) * fc;
  return mk3(first, first, first) + s * (1.f - fc);
}
__device__ __forceinline__ void basis(f3 w, f3& u, f3& v) {
  f3 a = (double)fabsf(w.x) > 0.1 ? mk3(0.f, 1.f, 0.f) : mk3(1.f, 0.f, 0.f);
  u = normalize(cross(a, w));
  v = cross(w, u);
  u = cross(v, w);
}
// (u, v, n): the tangent frame basis(n) gives.
__device__ __forceinline__ f3 sample_ggx(float xi0, float xi1, float rough, f3 n, f3 u, f3 v) {
  float a = rough * rough;
  float phi = (float)(2.0 * kPiD * (double)xi0);
  float ct = sqrtf((1.f - xi1) / (1.f + (a * a - 1.f) * xi1));
  float st = sqrtf(1.f - ct * ct);
  float sp, cp;
  sincos_def(phi, sp, cp);
  f3 h = mk3(st * cp, st * sp, ct);
  return (u * h.x + v * h.y) + n * h.z;
}
__device__ __forceinline__ f3 lambert(f3 diff) { return diff * (float)(1.0 / kPiD); }
// GetShadingResult (src/cudaRTBDPTStream.cu:642-663)
__device__ __forceinline__ f3 shading(f3 outDir, f3 inDir, f3 irrad, f3 n, f3 diff, float metal, float rough,
                                      float spec, float dw, float sw) {
  if (dot(n, inDir) <= 0.f) return mk3(0.f, 0.f, 0.f);
  f3 h = normalize(outDir + inDir);
  float voh = dot(outDir, h), nov = dot(n, outDir), noh = dot(n, h), nol = dot(n, inDir);
  f3 f0 = vlerp(mk3(1.f, 1.f, 1.f) * (0.08f * spec), diff, metal);
  f3 F = f_schlick(f0, voh);
  float dd = d_ggx(rough, noh), vv = vis_smith(rough, nov, nol);
  f3 specI = F * (dd * vv);
  f3 diffI = vmul(mk3(1.f, 1.f, 1.f) - F, lambert(diff));
  return vmul(irrad * nol, specI * sw + diffI * dw);
}

struct Surf {
  f3 norm, pos, diff, emissive;
  float trans, spec, metal, rough;
};
// tex2D of a float4 texture with normalized coordinates, wrap addressing and
// linear filtering (the reference's texture objects, src/cudaRTCommon.cu:416-441):
// the coordinate wraps to [0, 1), scales to texels, and the four texels around
// x - 0.5, y - 0.5 (indices wrapped) are blended in x, then in y.
__device__ __forceinline__ float4 tex_sample(const DevScene& sc, uint32_t id, float u, float v) {
  const uint4 d = sc.tex_desc[id];
  const int W = (int)d.y, H = (int)d.z;
  if (!(fabsf(u) < 1e30f)) u = 0.f;                 // NaN and infinities sample texel (0, 0)
  if (!(fabsf(v) < 1e30f)) v = 0.f;
  const float fu = u - floorf(u), fv = v - floorf(v);
  const float x = fu * (float)W - 0.5f, y = fv * (float)H - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float a = x - fx, b = y - fy;
  int i0 = min(max((int)fx, -1), W - 1), j0 = min(max((int)fy, -1), H - 1);
  if (i0 < 0) i0 += W;
  if (j0 < 0) j0 += H;
  const int i1 = i0 + 1 == W ? 0 : i0 + 1, j1 = j0 + 1 == H ? 0 : j0 + 1;
  const float4 t00 = sc.tex[d.x + (uint32_t)(j0 * W + i0)], t10 = sc.tex[d.x + (uint32_t)(j0 * W + i1)];
  const float4 t01 = sc.tex[d.x + (uint32_t)(j1 * W + i0)], t11 = sc.tex[d.x + (uint32_t)(j1 * W + i1)];
  const float ia = 1.f - a, ib = 1.f - b;
  float4 r;
  r.x = (t00.x * ia + t10.x * a) * ib + (t01.x * ia + t11.x * a) * b;
  r.y = (t00.y * ia + t10.y * a) * ib + (t01.y * ia + t11.y * a) * b;
  r.z = (t00.z * ia + t10.z * a) * ib + (t01.z * ia + t11.z * a) * b;
  r.w = (t00.w * ia + t10.w * a) * ib + (t01.w * ia + t11.w * a) * b;
  return r;
}

// Shading inputs at a hit: interpolated normal (and position), material, and
// GetMaterialColors' textures (src/cudaRTCommon.cu:369-406): a diffuse texture
// gives the colour and scales transparency by its alpha, an emissive texture
// gives the emission; texture ids outside the scene's textures mean none.
// The hit's surface from its triangle's 128-byte shading record: normals and
// material in one round trip after the hit; texture u and the vertex positions
// only for textured materials and first-hit G-buffers.
// Radiance the environment sends along direction d (no reference counterpart:
// the reference lights scenes with emissive triangles only).  Latitude-longitude
// map: u = atan2(d.z, d.x) / 2pi + 1/2, v = acos(d.y) / pi, texel centres at
// half-integers, bilinear between the four around (u W - 1/2, v H - 1/2), u
// wrapping and v clamped to the first and last rows.  A non-finite d sees 0.
__device__ __forceinline__ f3 env_radiance(const DevScene& sc, f3 d) {
  const int W = (int)sc.env_w, H = (int)sc.env_h;
  const float u = atan2_def(d.z, d.x) * (float)(0.5 / kPiD) + 0.5f;
  const float v = acos_def(fminf(fmaxf(d.y, -1.f), 1.f)) * (float)(1.0 / kPiD);
  const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
  if (!(fabsf(x) < 1e30f) || !(fabsf(y) < 1e30f)) return mk3(0.f, 0.f, 0.f);
  const float fx = floorf(x), fy = floorf(y);
  const float a = x - fx, b = y - fy;
  int i0 = (int)fx % W;
  if (i0 < 0) i0 += W;
  const int i1 = i0 + 1 == W ? 0 : i0 + 1;
  const int j0 = min(max((int)fy, 0), H - 1), j1 = min(max((int)fy + 1, 0), H - 1);
  const float4 t00 = sc.env[j0 * W + i0], t10 = sc.env[j0 * W + i1];
  const float4 t01 = sc.env[j1 * W + i0], t11 = sc.env[j1 * W + i1];
  const float ia = 1.f - a, ib = 1.f - b;
  return mk3((t00.x * ia + t10.x * a) * ib + (t01.x * ia + t11.x * a) * b,
             (t00.y * ia + t10.y * a) * ib + (t01.y * ia + t11.y * a) * b,
             (t00.z * ia + t10.z * a) * ib + (t01.z * ia + t11.z * a) * b);
}

__device__ __forceinline__ Surf fetch_surf(const DevScene& sc, uint32_t leaf, float w, float u, float v, bool want_pos) {
  Surf s;
  const float4* R = sc.surf + 8 * (size_t)leaf;
  const float4 r0 = R[0], r1 = R[1], r2 = R[2], md = R[3], me = R[4], mx = R[5];
  f3 n0 = ld3(r0), n1 = ld3(r1), n2 = ld3(r2);
  s.norm = (n0 * w + n1 * u) + n2 * v;
  const uint32_t dtid = __float_as_uint(mx.z), etid = __float_as_uint(mx.w);
  const bool dtex = dtid < sc.n_tex, etex = etid < sc.n_tex;
  s.diff = ld3(md);
  s.trans = md.w;
  s.emissive = ld3(me);
  if (want_pos) {
    const float4 ix = R[7];
    const float4 p0 = sc.vpos[__float_as_uint(ix.x)], p1 = sc.vpos[__float_as_uint(ix.y)],
                 p2 = sc.vpos[__float_as_uint(ix.z)];
    s.pos = (ld3(p0) * w + ld3(p1) * u) + ld3(p2) * v;
  } else {
    s.pos = mk3(0.f, 0.f, 0.f);
  }
  if (dtex || etex) {
    // uv = uv0 * w + uv1 * u + uv2 * v (cudaRTBDPTStream.cu:379-382)
    const float4 tu4 = R[6];
    const float tu = (tu4.x * w + tu4.y * u) + tu4.z * v, tvv = (r0.w * w + r1.w * u) + r2.w * v;
    if (dtex) {
      const float4 t = tex_sample(sc, dtid, tu, tvv);
      s.diff = mk3(t.x, t.y, t.z);
      s.trans = t.w * md.w;
    }
    if (etex) {
      const float4 t = tex_sample(sc, etid, tu, tvv);
      s.emissive = mk3(t.x, t.y, t.z);
    }
  }
  s.spec = me.w; s.metal = mx.x; s.rough = mx.y;
  return s;
}

enum { RT_EYE = 0, RT_DIFF = 1, RT_SPEC = 2, RT_LIGHT = 3 };
struct Bounce {
  f3 orig, dir, mul;
  int type;
};
// BRDF sample of one bounce (src/cudaRTBDPTStream.cu:228-329 and :402-502).
// tir_reflect: a refraction past the critical angle turns into a reflection
// (the path tracers, cudaRTPTStream.cu:152-160); the BDPT files take the
// refracted direction as computed.
__device__ Bounce sample_bounce(uint64_t& st, uint64_t inc, f3 ro, f3 rd, float t, const Surf& s, f3 nl, int cur,
                                bool refl_offset = false, bool tir_reflect = false) {
  float x0 = uniform01(st, inc), x1 = uniform01(st, inc);
  f3 u, v;                                           // one tangent frame for both samplers
  basis(nl, u, v);
  f3 hDir = sample_ggx(x0, x1, s.rough, nl, u, v);
  f3 ndir = rd * -1.f;
  float voH = dot(ndir, hDir);
  f3 f0 = vlerp(mk3(0.08f * s.spec, 0.08f * s.spec, 0.08f * s.spec), s.diff, s.metal);
  f3 F = f_schlick(f0, voH);
  float reflProb = lerpf(length(F), 1.f, s.metal);
  float refrProb = s.trans;
  f3 reflDir = mk3(0.f, 0.f, 0.f), refrDir = mk3(0.f, 0.f, 0.f);
  Bounce b;
  b.orig = ro; b.dir = rd; b.type = cur; b.mul = mk3(0.f, 0.f, 0.f);
  if (refrProb > 0.f) {
    bool into = dot(nl, s.norm) > 0.f;
    float nt = s.spec * 0.8f + 1.f, nc = 1.f;
    float nnt = into ? nc / nt : nt / nc;
    float ddn = dot(hDir, rd);
    float cos2t = 1.f - nnt * nnt * (1.f - ddn * ddn);
    if (tir_reflect && cos2t < 0.f) reflProb = 1.f;
    else refrDir = normalize(rd * nnt - hDir * (ddn * nnt + sqrtf(cos2t)));
  }
  if (reflProb > 0.f) {
    reflDir = normalize(rd - (hDir * 2.f) * dot(hDir, rd));
    if (dot(reflDir, nl) < 0.f) reflProb = 0.f;
  }
  if (prob_rand(st, inc, reflProb)) {
    b.orig = ro + rd * (float)((double)t - kBiasEps);
    // the adaptive files also step off along the reflection (AdapBilatAtomic.cu:584)
    if (refl_offset) b.orig = b.orig + reflDir * (float)kBiasEps;
    b.dir = reflDir;
    float VoH = dot(ndir, hDir), NoV = dot(nl, ndir), NoH = dot(nl, hDir), NoL = dot(nl, reflDir);
    float G = vis_smith(s.rough, NoV, NoL);
    b.mul = (F * G * VoH) / (NoV * NoH * reflProb);
    b.type = RT_SPEC;
  } else if (prob_rand(st, inc, refrProb)) {
    b.orig = (ro + rd * (float)((double)t + kBiasEps)) + refrDir * (float)kBiasEps;
    b.dir = refrDir;
    float cosine = dot(nl * -1.f, refrDir);
    b.mul = s.diff * cosine / (refrProb * (1.f - reflProb));
    b.type = RT_SPEC;
  } else {
    float r1 = (float)(2.0 * kPiD * (double)uniform01(st, inc));
    float r2cos = sqrtf(uniform01(st, inc));
    float r2sin = 1.f - r2cos * r2cos;
    float s1, c1;
    sincos_def(r1, s1, c1);
    f3 dd = normalize((nl * r2cos + (u * r2sin) * c1) + (v * r2sin) * s1);
    b.orig = (ro + rd * t) + dd * (float)kBiasEps;
    b.dir = dd;
    b.mul = lambert(s.diff) * (float)kPiD / ((1.f - refrProb) * (1.f - reflProb));
    b.type = RT_DIFF;
  }
  return b;
}

__device__ __forceinline__ double shadow_max(double dist, uint32_t flags) {
  return (flags & 1u) ? dist - kBiasEps : dist - 2.0 * kBiasEps;
}

// First frame of the batch: the argument, or the device counter of a replayed graph.
__device__ __forceinline__ uint32_t frame_base(const FrameArgs& fa) {
  return fa.frame_dev ? *fa.frame_dev : fa.frame0;
}

// View of a batch-global pixel (0 without a view table).
__device__ __forceinline__ uint32_t view_of(const FrameArgs& fa, uint32_t pix) {
  uint32_t lo = 0, hi = fa.n_views;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (fa.views[mid].offset <= pix) lo = mid; else hi = mid;
  }
  return lo;
}
// The pixel's index within its view: the PCG32 stream of an eye path is keyed
// by it, so every view of a batch traces the samples a one-view renderer would.
__device__ __forceinline__ uint32_t local_pixel(const FrameArgs& fa, uint32_t pix) {
  return fa.views ? pix - fa.views[view_of(fa, pix)].offset : pix;
}

// ---------------------------------------------------------------- frame
__global__ void k_fill_u32(uint32_t* p, uint32_t v) { *p = v; }
__global__ void k_fill_n_u32(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}
__global__ void k_frame_begin(FrameCounters* fc, uint32_t n_paths) {
  uint32_t i = threadIdx.x;
  if (i < 16) { fc->active[i] = i == 0 ? n_paths : 0u; fc->shadow[i] = 0u; }
}
hipError_t launch_frame_begin(FrameCounters* fc, uint32_t n_paths, hipStream_t st) {
  k_frame_begin<<<1, 64, 0, st>>>(fc, n_paths);
  return hipGetLastError();
}
__global__ void k_frame_begin_dev(FrameCounters* fc, const uint32_t* total, uint32_t cap) {
  uint32_t i = threadIdx.x;
  uint32_t t = *total;
  if (i < 16) { fc->active[i] = i == 0 ? (t < cap ? t : cap) : 0u; fc->shadow[i] = 0u; }
}
hipError_t launch_frame_begin_dev(FrameCounters* fc, const uint32_t* total, uint32_t cap, hipStream_t st) {
  k_frame_begin_dev<<<1, 64, 0, st>>>(fc, total, cap);
  return hipGetLastError();
}

__global__ void k_frame_end(FrameCounters* fc, DevStats* stats, uint32_t nframes, uint32_t depth) {
  if (threadIdx.x == 0) {
    unsigned long long e = 0, s = 0;
    for (uint32_t b = 0; b < depth && b < 16; b++) { e += fc->active[b]; s += fc->shadow[b]; }
    stats->ext_rays += e;
    stats->shadow_rays += s;
    stats->eye_paths += fc->active[0];
    stats->frames += nframes;
  }
}
hipError_t launch_frame_end(FrameCounters* fc, DevStats* stats, uint32_t nframes, uint32_t depth, hipStream_t st) {
  k_frame_end<<<1, 64, 0, st>>>(fc, stats, nframes, depth);
  return hipGetLastError();
}
__global__ void k_frame_advance(uint32_t* frame_dev, uint32_t k) {
  if (threadIdx.x == 0) *frame_dev += k;
}
hipError_t launch_frame_advance(uint32_t* frame_dev, uint32_t k, hipStream_t st) {
  if (!frame_dev) return hipErrorInvalidValue;
  k_frame_advance<<<1, 64, 0, st>>>(frame_dev, k);
  return hipGetLastError();
}

__device__ __forceinline__ void store_lv(LightVert& V, f3 pos, f3 norm, f3 irrad, f3 irradDir, const Surf& s) {
  V.pos[0] = pos.x; V.pos[1] = pos.y; V.pos[2] = pos.z;
  V.norm[0] = norm.x; V.norm[1] = norm.y; V.norm[2] = norm.z;
  V.irrad[0] = irrad.x; V.irrad[1] = irrad.y; V.irrad[2] = irrad.z;
  V.irradDir[0] = irradDir.x; V.irradDir[1] = irradDir.y; V.irradDir[2] = irradDir.z;
  V.diff[0] = s.diff.x; V.diff[1] = s.diff.y; V.diff[2] = s.diff.z;
  V.emissive[0] = s.emissive.x; V.emissive[1] = s.emissive.y; V.emissive[2] = s.emissive.z;
  V.specular = s.spec; V.metallic = s.metal; V.roughness = s.rough;
  V.pad[0] = V.pad[1] = V.pad[2] = 0.f;
}

// Light-vertex caches, one workgroup per cache slot to build (light_paths <= kBlock).
__global__ __launch_bounds__(kBlock) void k_light_pass(DevScene sc, FrameArgs fa, LightJobs jobs,
                                                       LightVert* __restrict__ lv_all,
                                                       FrameCounters* __restrict__ fc) {
  __shared__ uint64_t lstk[CPT_LDS_STACK * kBlock];
  __shared__ uint32_t wcnt[kBlock / 64];
  const uint32_t x = threadIdx.x, L = fa.light_paths;
  uint32_t slot, job_frame;
  if (jobs.dev_groups) {                       // groups of the device-counted batch
    const uint32_t f0 = frame_base(fa), g = f0 / 3u + blockIdx.x;
    if (g > (f0 + fa.nframes - 1u) / 3u) return;
    slot = g % (uint32_t)kLightSlots;
    job_frame = 3u * g;
  } else {
    slot = jobs.slot[blockIdx.x];
    job_frame = jobs.frame[blockIdx.x];
  }
  LightVert* __restrict__ lv = lv_all + (size_t)slot * fa.lv_stride;
  const uint64_t light_key = stream_key(fa.seed, job_frame, 2u, 0u);
  const uint64_t inc = ((uint64_t)x << 1) | 1u;
  uint64_t st = 0;
  bool live = false;
  f3 orig = mk3(0.f, 0.f, 0.f), dir = orig, smp = orig, mul = mk3(1.f, 1.f, 1.f);
  int type = RT_LIGHT;
  uint32_t depth = 0;
  if (sc.n_light_tris == 0) {
    if (x == 0) fc->light_n[slot] = 0;
    return;
  }
  if (x < L) {
    pcg32_seed(st, inc, light_key);
    uint32_t id = (uint32_t)(uniform01(st, inc) * (float)sc.n_light_tris);
    if (id >= sc.n_light_tris) id = sc.n_light_tris - 1;
    float lw = uniform01(st, inc), lu = uniform01(st, inc);
    if (lw + lu > 1.f) { lw = 1.f - lw; lu = 1.f - lu; }
    float lvv = (1.f - lw) - lu;
    Surf s = fetch_surf(sc, sc.light_tris[id], lw, lu, lvv, true);
    f3 u, v;
    basis(s.norm, u, v);
    float r1 = (float)(2.0 * kPiD * (double)uniform01(st, inc));
    float r2cos = sqrtf(uniform01(st, inc));
    float r2sin = 1.f - r2cos * r2cos;
    float s1, c1;
    sincos_def(r1, s1, c1);
    f3 dd = normalize((s.norm * r2cos + (u * r2sin) * c1) + (v * r2sin) * s1);
    orig = s.pos + s.norm * (float)kBiasEps;
    dir = dd;
    smp = s.emissive;
    live = true;
    store_lv(lv[x], s.pos, s.norm, s.emissive, mk3(0.f, 0.f, 0.f), s);
  }
  __syncthreads();
  uint32_t cur = L;
  const int lane = x & 63, wid = x >> 6;
  for (;;) {
    unsigned long long m = __ballot(live);
    uint32_t below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[wid] = __popcll(m);
    __syncthreads();
    uint32_t pre = 0, total = 0;
    for (int w = 0; w < kBlock / 64; w++) { if (w < wid) pre += wcnt[w]; total += wcnt[w]; }
    __syncthreads();
    if (total == 0) break;
    if (live) {
      uint32_t slot = cur + pre + below;
      TraceHit h;
      uint32_t nn = 0, nt = 0;
      if (!traverse4<false>(sc, orig, dir, kInfF, (float)kBiasEps, false, (LdsU64*)(lstk + x), h, nn, nt)) {
        lv[slot] = lv[x];
        live = false;
      } else {
        float w = (1.f - h.u) - h.v;
        Surf s = fetch_surf(sc, h.leaf, w, h.u, h.v, true);
        f3 nl = dot(s.norm, dir) < 0.f ? s.norm : s.norm * -1.f;
        Bounce b = sample_bounce(st, inc, orig, dir, h.t, s, nl, type);
        f3 irr = b.type != RT_DIFF ? mk3(0.f, 0.f, 0.f) : smp;
        f3 vn = dot(b.dir, nl) < 0.f ? nl * -1.f : nl;
        store_lv(lv[slot], s.pos, vn, irr, dir * -1.f, s);
        smp = s.emissive + vmul(smp, b.mul);
        float contrib = length(mul) * length(b.mul);
        if (length(s.emissive) > 0.f) contrib = 0.f;
        if (uniform01(st, inc) > contrib || depth + 1 >= fa.light_depth) {
          live = false;
        } else {
          mul = vmul(mul, b.mul);
          dir = b.dir;
          orig = b.orig;
          type = b.type;
          depth++;
        }
      }
    }
    cur += total;
    __syncthreads();
  }
  if (x == 0) fc->light_n[slot] = cur;
}
hipError_t launch_light_pass(const DevScene& sc, const FrameArgs& fa, const LightJobs& jobs, LightVert* lv,
                             FrameCounters* fc, hipStream_t st) {
  if (jobs.n == 0 && jobs.dev_groups == 0) return hipSuccess;
  if (!lv || !fc || fa.light_paths > (uint32_t)kBlock || (sc.n_light_tris && !sc.light_tris) ||
      jobs.n > (uint32_t)kLightSlots || jobs.dev_groups > (uint32_t)kLightSlots ||
      (jobs.dev_groups && !fa.frame_dev))
    return hipErrorInvalidValue;
  for (uint32_t j = 0; j < jobs.n; j++)
    if (jobs.slot[j] >= (uint32_t)kLightSlots) return hipErrorInvalidValue;
  k_light_pass<<<jobs.dev_groups ? jobs.dev_groups : jobs.n, kBlock, 0, st>>>(sc, fa, jobs, lv, fc);
  return hipGetLastError();
}

// Running-mean update of pt_applyPathQueueResult_kernel
// (src/cudaRTBDPTStream.cu:754-776); one path per pixel per frame.
__device__ __forceinline__ void apply_pixel(float* result, uint32_t* sample_n, uint32_t pix, uint32_t frame, f3 acc,
                                            uint32_t n) {
  uint32_t nold = frame == 0 ? 0u : sample_n[pix];
  if (frame == 0) sample_n[pix] = 0u;
  uint32_t nn = nold + n;
  if (!(nn > nold) || isinf3(acc)) return;
  float ri = 1.f / (float)nn, oi = (float)nold * ri;
  float* r = result + (size_t)pix * 3;
  r[0] = fmaxf(ri * acc.x + oi * r[0], 0.f);
  r[1] = fmaxf(ri * acc.y + oi * r[1], 0.f);
  r[2] = fmaxf(ri * acc.z + oi * r[2], 0.f);
  sample_n[pix] = nn;
}

// One eye path in registers: the ray (o, max distance; d, epsilon), pixel,
// sample slot, depth, frame offset in the batch, throughput, emission gathered
// so far, the accumulated radiance with its sample count (acc.w, as bits) and
// the RNG state.  PathBuf holds the same fields as columns.
struct PathState {
  float4 ro, rd, mul, smp, acc;
  uint32_t pix, sid, depth, fk;
  uint32_t lim;    // depth bound of this path (0: the frame's eye depth)
  uint32_t type;   // type of the ray being traced (RT_EYE, RT_DIFF, RT_SPEC)
  uint64_t rng;
};
// PathBuf::depth packs depth (bits 0-7), frame offset (8-15), depth bound
// (16-23), ray type (24-25) and kFreshPath: a path k_gen_eye started, whose
// mul / smp / acc columns were not written (throughput 1, nothing gathered).
constexpr uint32_t kFreshPath = 1u << 26;
__device__ __forceinline__ uint32_t pack_dk(const PathState& p) {
  return p.depth | (p.fk << 8) | (p.lim << 16) | (p.type << 24);
}
__device__ __forceinline__ void unpack_dk(PathState& p, uint32_t dk) {
  p.depth = dk & 0xFFu;
  p.fk = (dk >> 8) & 0xFFu;
  p.lim = (dk >> 16) & 0xFFu;
  p.type = (dk >> 24) & 3u;
}
// The connection's shadow ray: origin and max distance, direction and epsilon,
// contribution added to the path's accumulator when nothing occludes it.
// A shadow ray's target slot with this bit set adds its contribution without
// the connection's N += 4 (the Disney mode's light samples are part of their
// path's one sample).  Slots stay below 2^31.
constexpr uint32_t kShadowNoCount = 1u << 31;
struct ShadowRay {
  float4 o, d, c;
};

// Eye ray of pixel (x, y) drawn from PCG32 stream `key`, pixel as sequence
// (pt_genPathQueue_kernel, src/cudaRTBDPTStream.cu:619-640; the adaptive queue
// of AdapBilatAtomic.cu:826-840 uses the same pinhole).  In a multi-view batch
// `pix` is the batch-global pixel: the camera and the pixel's position come
// from its view, the RNG stream stays keyed by the global pixel.
// st: the path's PCG32 state, positioned where the jitter draws start.
__device__ __forceinline__ PathState eye_path_at(const FrameArgs& fa, uint32_t pix, uint32_t sid, uint64_t st) {
  uint32_t lp = pix, vw = fa.width, vh = fa.height;
  const float *cpos = fa.cam_pos, *cdir = fa.cam_dir, *cup = fa.cam_up, *cright = fa.cam_right;
  float tf = fa.tan_half;
  if (fa.views) {
    const ViewDesc& V = fa.views[view_of(fa, pix)];
    lp = pix - V.offset;
    vw = V.width;
    vh = V.height;
    cpos = V.cam_pos; cdir = V.cam_dir; cup = V.cam_up; cright = V.cam_right;
    tf = V.tan_half;
  }
  uint32_t y = lp / vw, x = lp - y * vw;
  float W = (float)vw, H = (float)vh;
  float u = ((((2.f * ((float)x + 0.5f)) / W - 1.f) * tf) * W) / H;
  float v = ((2.f * ((float)y + 0.5f)) / H - 1.f) * tf;
  const uint64_t inc = ((uint64_t)lp << 1) | 1u;
  float au = u + ((uniform01(st, inc) - 0.5f) / H) * tf;
  float av = v + ((uniform01(st, inc) - 0.5f) / H) * tf;
  f3 right = mk3(cright[0], cright[1], cright[2]);
  f3 up = mk3(cup[0], cup[1], cup[2]);
  f3 cd = mk3(cdir[0], cdir[1], cdir[2]);
  f3 d = normalize((right * au + up * av) + cd);
  PathState p;
  p.ro = make_float4(cpos[0], cpos[1], cpos[2], kInfF);
  p.rd = make_float4(d.x, d.y, d.z, (float)kBiasEps);
  p.pix = pix;
  p.sid = sid;
  p.depth = 0;
  p.fk = 0;
  p.lim = 0;
  p.type = RT_EYE;
  p.mul = make_float4(1.f, 1.f, 1.f, 0.f);
  p.smp = make_float4(0.f, 0.f, 0.f, 0.f);
  p.acc = make_float4(0.f, 0.f, 0.f, 0.f);
  p.rng = st;
  return p;
}
__device__ __forceinline__ PathState eye_path(const FrameArgs& fa, uint32_t pix, uint32_t sid, uint64_t key) {
  const uint64_t inc = ((uint64_t)local_pixel(fa, pix) << 1) | 1u;
  uint64_t st;
  pcg32_seed(st, inc, key);
  return eye_path_at(fa, pix, sid, st);
}
__device__ __forceinline__ void write_eye_path(const FrameArgs& fa, PathBuf& pb, uint32_t slot, uint32_t pix,
                                               uint32_t sid, uint64_t key) {
  const PathState p = eye_path(fa, pix, sid, key);
  pb.ray_o[slot] = p.ro;
  pb.ray_d[slot] = p.rd;
  pb.pix[slot] = p.pix;
  pb.sid[slot] = p.sid;
  pb.depth[slot] = 0;
  pb.mul[slot] = p.mul;
  pb.smp[slot] = p.smp;
  pb.acc[slot] = p.acc;
  pb.rng[slot] = p.rng;
}

__device__ __forceinline__ void zero_rec(SampleRec& rec, uint32_t q) {
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  rec.acc[q] = z; rec.gpos[q] = z; rec.gnrm[q] = z; rec.gdif[q] = z;
}

// One path per pixel and frame of the batch.  Path slots are pixel-major
// (slot = pixel * frames + frame offset): a wave's first rays are one pixel's
// jittered rays of consecutive frames, which visit the same nodes, where
// frame-major slots give it 64 neighbouring pixels (extension 21 % faster on
// the 262k-triangle grid at 64 frames).  The sample record stays
// frame-major (sid = frame offset * W*H + pixel) for k_apply_frames.
// The eye path of slot i and its packed depth word.  The first extension and
// shading of a batch can call this in place of reading what k_gen_eye wrote.
__device__ __forceinline__ PathState eye_slot(const FrameArgs& fa, uint32_t i, uint32_t& dk) {
  const uint32_t wh = fa.width * fa.height, K = fa.nframes, pix = i / K, k = i - pix * K;
  dk = (k << 8) | kFreshPath;
  return eye_path(fa, pix, k * wh + pix, stream_key(fa.seed, frame_base(fa) + k, 1u, 0u));
}
__global__ __launch_bounds__(kBlock) void k_gen_eye(FrameArgs fa, PathBuf pb, SampleRec rec) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= fa.width * fa.height * fa.nframes) return;
  uint32_t dk;
  const PathState p = eye_slot(fa, i, dk);
  pb.ray_o[i] = p.ro;
  pb.ray_d[i] = p.rd;
  pb.pix[i] = p.pix;
  pb.sid[i] = p.sid;
  pb.depth[i] = dk;
  pb.rng[i] = p.rng;
  if (rec.acc) zero_rec(rec, p.sid);
}
hipError_t launch_gen_eye(const FrameArgs& fa, PathBuf pb, hipStream_t st) {
  if (!pathbuf_ok(pb)) return hipErrorInvalidValue;
  uint32_t n = fa.width * fa.height * fa.nframes;
  SampleRec none = {nullptr, nullptr, nullptr, nullptr};
  k_gen_eye<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(fa, pb, none);
  return hipGetLastError();
}
hipError_t launch_gen_eye_rec(const FrameArgs& fa, PathBuf pb, SampleRec rec, hipStream_t st) {
  if (!pathbuf_ok(pb) || !rec_ok(rec)) return hipErrorInvalidValue;
  uint32_t n = fa.width * fa.height * fa.nframes;
  k_gen_eye<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(fa, pb, rec);
  return hipGetLastError();
}

// The path tracers' running mean (pt_applyPathQueueResult_kernel,
// src/cudaRTPTStream.cu:303-325): the frame's per-pixel mean acc / n weighted
// 1 / (frame + 1), whatever the number of samples the pixel took.
__device__ __forceinline__ void apply_pixel_pt(float* result, uint32_t* sample_n, uint32_t pix, uint32_t frame, f3 acc,
                                               uint32_t n) {
  const uint32_t nold = frame == 0 ? 0u : sample_n[pix];
  if (n > 0) {
    const float fn = (float)n;
    const f3 sr = mk3(acc.x / fn, acc.y / fn, acc.z / fn);
    const float ri = 1.f / (float)(frame + 1), oi = 1.f - ri;
    float* r = result + (size_t)pix * 3;
    r[0] = fmaxf(ri * sr.x + oi * r[0], 0.f);
    r[1] = fmaxf(ri * sr.y + oi * r[1], 0.f);
    r[2] = fmaxf(ri * sr.z + oi * r[2], 0.f);
  }
  sample_n[pix] = nold + n;
}

// Frames of the batch applied in order: pt_applyPathQueueResult_kernel's
// running mean (src/cudaRTBDPTStream.cu:741-778) once per frame (pt: the path
// tracers' mean).
__global__ __launch_bounds__(kBlock) void k_apply_frames(FrameArgs fa, const float4* __restrict__ rec_acc,
                                                         float* __restrict__ result, uint32_t* __restrict__ sample_n,
                                                         int pt) {
  const uint32_t wh = fa.width * fa.height;
  const uint32_t pix = blockIdx.x * kBlock + threadIdx.x;
  if (pix >= wh) return;
  const uint32_t f0 = frame_base(fa);
  for (uint32_t k = 0; k < fa.nframes; k++) {
    float4 a = rec_acc[(size_t)k * wh + pix];
    if (pt) apply_pixel_pt(result, sample_n, pix, f0 + k, ld3(a), __float_as_uint(a.w));
    else apply_pixel(result, sample_n, pix, f0 + k, ld3(a), __float_as_uint(a.w));
  }
}
hipError_t launch_apply_frames(const FrameArgs& fa, const float4* rec_acc, float* result, uint32_t* sample_n,
                               int pt, hipStream_t st) {
  if (!rec_acc || !result || !sample_n) return hipErrorInvalidValue;
  uint32_t wh = fa.width * fa.height;
  k_apply_frames<<<(wh + kBlock - 1) / kBlock, kBlock, 0, st>>>(fa, rec_acc, result, sample_n, pt);
  return hipGetLastError();
}

hipError_t launch_extend(const DevScene& sc, PathBuf pb, const FrameCounters* fc, int bounce, uint32_t max_n,
                         float4* hits, DevStats* stats, int lds_stack, int quant, hipStream_t st) {
  if (max_n == 0) return hipSuccess;
  if (!pathbuf_ok(pb) || !fc || !hits || !sc.tris || !(quant ? (const void*)sc.nodes4q : (const void*)sc.nodes4))
    return hipErrorInvalidValue;
  dim3 grid((max_n + kBlock - 1) / kBlock), block(kBlock);
  const uint32_t* cnt = &fc->active[bounce];
#define CPT_EXT(L, S, W) k_trace<L, false, 1, S, W><<<grid, block, 0, st>>>(sc, pb.ray_o, pb.ray_d, 1, cnt, max_n, 0, hits, stats)
  if (quant) {
#define CPT_EXT_Q(S, W) CPT_EXT(2, S, W)
#define CPT_EXT_F(S, W) CPT_EXT(0, S, W)
    CPT_VARIANT_SWITCH(lds_stack, CPT_EXT_Q)
  } else {
    CPT_VARIANT_SWITCH(lds_stack, CPT_EXT_F)
  }
#undef CPT_EXT_Q
#undef CPT_EXT_F
#undef CPT_EXT
  return hipGetLastError();
}

// Persistent closest-hit grid over a global queue of the bounce's live rays.
template <int STK, int MINW, bool Q>
__global__ __launch_bounds__(kBlock, MINW) void k_extend_persist(DevScene sc, const float4* __restrict__ ro,
                                                                 const float4* __restrict__ rd,
                                                                 const uint32_t* __restrict__ countp, uint32_t chunk,
                                                                 uint32_t* __restrict__ queue,
                                                                 float4* __restrict__ out,
                                                                 DevStats* __restrict__ stats) {
  __shared__ uint64_t lstk[STK * kBlock];
  __shared__ uint32_t base_s;
  const uint32_t count = *countp;
  uint32_t nn = 0, nt = 0;
  for (;;) {
    if (threadIdx.x == 0) base_s = atomicAdd(queue, chunk);
    __syncthreads();
    const uint32_t base = base_s;
    __syncthreads();
    if (base >= count) break;
    const uint32_t end = min(base + chunk, count);
    for (uint32_t i = base + threadIdx.x; i < end; i += kBlock) {
      float4 a = ro[i], b = rd[i];
      TraceHit h;
      bool hit = traverse4<false, STK, Q>(sc, ld3(a), ld3(b), a.w, b.w, false, lds_column(lstk), h, nn, nt);
      float4 r;
      r.x = hit ? h.t : kInfF;
      r.y = __uint_as_float(hit ? h.leaf : kInvalid);
      r.z = h.u;
      r.w = h.v;
      out[i] = r;
    }
  }
  if (stats) {
    uint32_t sn = wave_sum(nn), st = wave_sum(nt);
    if ((threadIdx.x & 63) == 0 && (sn | st)) {
      atomicAdd(&stats->ext_nodes, (unsigned long long)sn);
      atomicAdd(&stats->ext_tris, (unsigned long long)st);
    }
  }
}
hipError_t launch_extend_persist(const DevScene& sc, PathBuf pb, const FrameCounters* fc, int bounce, float4* hits,
                                 DevStats* stats, int lds_stack, int quant, uint32_t chunk, uint32_t grid,
                                 uint32_t* queue, hipStream_t st) {
  if (!pathbuf_ok(pb) || !fc || !hits || !sc.tris || !queue || chunk == 0 || grid == 0 ||
      !(quant ? (const void*)sc.nodes4q : (const void*)sc.nodes4))
    return hipErrorInvalidValue;
  k_fill_u32<<<1, 1, 0, st>>>(queue, 0u);
  const uint32_t* cnt = &fc->active[bounce];
#define CPT_EXTP(S, W, Q) k_extend_persist<S, W, Q><<<grid, kBlock, 0, st>>>(sc, pb.ray_o, pb.ray_d, cnt, chunk, queue, hits, stats)
#define CPT_EXTP_Q(S, W) CPT_EXTP(S, W, true)
#define CPT_EXTP_F(S, W) CPT_EXTP(S, W, false)
  if (quant) {
    CPT_VARIANT_SWITCH(lds_stack, CPT_EXTP_Q)
  } else {
    CPT_VARIANT_SWITCH(lds_stack, CPT_EXTP_F)
  }
#undef CPT_EXTP_Q
#undef CPT_EXTP_F
#undef CPT_EXTP
  return hipGetLastError();
}

// Dynamic refill: a persistent grid whose lanes each trace one ray at a time
// and take the next as soon as theirs is finished.  A wave hands out ray
// indices from a chunk it took from the global counter `queue` (one atomic per
// chunk): kDynChunk rays when the launch has enough for every wave to take
// several, else 64, so that a small batch still spreads over every wave; lanes
// refill once at least `refill` (1..64) are idle.
// A finished lane writes its result at the wave's next refill point, together
// with the other finished lanes, not in the traversal loop: the shadow
// epilogue's read-modify-write would otherwise stall the whole wave on its load
// every time one lane finishes.
// OUT 1: closest hit (t, leaf, u, v) into out[ray]; OUT 3: the connection's
// visibility (k_shadow's epilogue).
constexpr uint32_t kDynChunk = 256;
// EYE (closest hit, quantized nodes): ray i is the eye ray of path slot i of
// the batch fa describes, computed in place of reading ro / rd.  A separate
// instantiation, so that the plain kernels' code is not touched by it.
template <int OUT, int STK, int MINW, bool Q, bool EYE = false>
__global__ __launch_bounds__(kBlock, MINW) void k_trace_dyn(DevScene sc, const float4* __restrict__ ro,
                                                            const float4* __restrict__ rd,
                                                            const uint32_t* __restrict__ countp,
                                                            uint32_t* __restrict__ queue, float4* __restrict__ out,
                                                            const float4* __restrict__ contrib, PathBuf nxt,
                                                            DevStats* __restrict__ stats, uint32_t refill,
                                                            FrameArgs fa) {
  constexpr bool ANY = OUT == 3;
  __shared__ uint64_t lstk_all[STK * kBlock];
  LdsU64* lstk = lds_column(lstk_all);
  uint64_t pstk[CPT_STACK_TOTAL - STK];
  const uint32_t count = *countp;
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t waves = gridDim.x * (kBlock / 64);
  const uint32_t chunk = (uint64_t)count >= (uint64_t)kDynChunk * waves ? kDynChunk : 64u;
  // When one chunk per wave covers the launch, wave w takes chunk w and no
  // atomic is issued (every wave's atomic on the one counter would cost more
  // than a small launch's traversal); otherwise every chunk comes from the
  // counter, in the order waves ask.
  const bool fixed = (uint64_t)count <= (uint64_t)waves * chunk;
  const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const uint64_t first = fixed ? (uint64_t)wave * chunk : (uint64_t)count;
  uint32_t cnext = (uint32_t)min<uint64_t>(first, count);   // the wave's chunk (uniform across the wave)
  uint32_t cend = (uint32_t)min<uint64_t>(first + chunk, count);
  bool exhausted = false;                  // uniform
  bool active = false;
  bool pending = false;                    // finished, result not yet written
  uint32_t ray = 0;
  RayState R;
  uint32_t nn = 0, nt = 0;
  __shared__ uint32_t spill_all[kBlock / 64];
  uint32_t* spill = spill_all + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (lane == 0) *spill = 0u;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= refill && pending) {
      pending = false;
      const bool hit = R.h.leaf != kInvalid;
      if (OUT == 1) {
        float4 r;
        r.x = hit ? R.h.t : kInfF;
        r.y = __uint_as_float(hit ? R.h.leaf : kInvalid);
        r.z = R.h.u;
        r.w = R.h.v;
        out[ray] = r;
      } else if (!hit) {
        const float4 c = contrib[ray];
        const uint32_t bits = __float_as_uint(c.w), slot = bits & ~kShadowNoCount;
        float4 acc = nxt.acc[slot];
        acc.x = acc.x + c.x;
        acc.y = acc.y + c.y;
        acc.z = acc.z + c.z;
        acc.w = __uint_as_float(__float_as_uint(acc.w) + ((bits & kShadowNoCount) ? 0u : 4u));
        nxt.acc[slot] = acc;
      }
    }
    if (!exhausted && n_idle >= refill) {
      const uint32_t rank = (uint32_t)__popcll(idle & below);
      uint32_t given = 0;
      while (given < n_idle) {
        if (cnext >= cend) {
          if (fixed) { exhausted = true; break; }
          uint32_t q = 0;
          if (lane == 0) q = atomicAdd(queue, chunk);
          const uint64_t b = __shfl(q, 0, 64);
          if (b >= count) { exhausted = true; break; }
          cnext = (uint32_t)b;
          cend = (uint32_t)min<uint64_t>(b + chunk, count);
        }
        const uint32_t take = min(n_idle - given, cend - cnext);
        if (!active && rank >= given && rank < given + take) {
          ray = cnext + (rank - given);
          if (EYE) {
            uint32_t dk;
            const PathState e = eye_slot(fa, ray, dk);
            ray_begin(R, ld3(e.ro), ld3(e.rd), e.ro.w, e.rd.w, sc.qmag);
          } else {
            const float4 a = ro[ray], b4 = rd[ray];
            ray_begin(R, ld3(a), ld3(b4), a.w, b4.w, sc.qmag);
          }
          active = true;
        }
        cnext += take;
        given += take;
      }
    }
    // every lane idle with nothing left to take: the results were written above
    if (__ballot(active) == 0ull) break;
    if (active && ray_step<ANY, STK, Q>(sc, R, false, lstk, pstk, nn, nt, spill)) {
      active = false;
      pending = true;
    }
  }
  if (stats) {
    uint32_t sn = wave_sum(nn), st = wave_sum(nt);
    const uint32_t ss = *spill;         // the wave's own word, after its last atomic
    if (lane == 0 && (sn | st)) {
      atomicAdd(OUT == 1 ? &stats->ext_nodes : &stats->sh_nodes, (unsigned long long)sn);
      atomicAdd(OUT == 1 ? &stats->ext_tris : &stats->sh_tris, (unsigned long long)st);
      if (ss) atomicAdd(OUT == 1 ? &stats->ext_spill : &stats->sh_spill, (unsigned long long)ss);
    }
  }
}
hipError_t launch_trace_dyn(const DevScene& sc, const float4* ro, const float4* rd, const uint32_t* count,
                            float4* out, const float4* contrib, PathBuf nxt, DevStats* stats, int lds_stack,
                            int quant, int shadow, uint32_t grid, uint32_t refill, uint32_t* queue,
                            hipStream_t st, const FrameArgs* eye) {
  if (!ro || !rd || !count || !queue || grid == 0 || refill < 1 || refill > 64 || !sc.tris ||
      !(quant ? (const void*)sc.nodes4q : (const void*)sc.nodes4))
    return hipErrorInvalidValue;
  if (shadow ? (!contrib || !pathbuf_ok(nxt) || eye) : !out) return hipErrorInvalidValue;
  const FrameArgs fa = eye ? *eye : FrameArgs{};
  k_fill_u32<<<1, 1, 0, st>>>(queue, 0u);
#define CPT_DYN(O, S, W, Qb) k_trace_dyn<O, S, W, Qb><<<grid, kBlock, 0, st>>>(sc, ro, rd, count, queue, out, contrib, nxt, stats, refill, fa)
#define CPT_DYN_1E(S, W) k_trace_dyn<1, S, W, true, true><<<grid, kBlock, 0, st>>>(sc, ro, rd, count, queue, out, contrib, nxt, stats, refill, fa)
#define CPT_DYN_3Q(S, W) CPT_DYN(3, S, W, true)
#define CPT_DYN_3F(S, W) CPT_DYN(3, S, W, false)
#define CPT_DYN_1Q(S, W) CPT_DYN(1, S, W, true)
#define CPT_DYN_1F(S, W) CPT_DYN(1, S, W, false)
  if (shadow) {
    if (quant) { CPT_VARIANT_SWITCH(lds_stack, CPT_DYN_3Q) } else { CPT_VARIANT_SWITCH(lds_stack, CPT_DYN_3F) }
  } else if (eye) {
    if (!quant) return hipErrorInvalidValue;
    CPT_VARIANT_SWITCH(lds_stack, CPT_DYN_1E)
  } else {
    if (quant) { CPT_VARIANT_SWITCH(lds_stack, CPT_DYN_1Q) } else { CPT_VARIANT_SWITCH(lds_stack, CPT_DYN_1F) }
  }
#undef CPT_DYN_3Q
#undef CPT_DYN_3F
#undef CPT_DYN_1Q
#undef CPT_DYN_1F
#undef CPT_DYN_1E
#undef CPT_DYN
  return hipGetLastError();
}

// Same plus the count-weighted first-hit G-buffer running means of the
// adaptive file (src/cudaRTBDPTStreamAdapBilatAtomic.cu:999-1036).
__device__ __forceinline__ void apply_pixel_gbuf(const Accum& ac, uint32_t pix, uint32_t frame, f3 acc, uint32_t n,
                                                 uint32_t sid) {
  uint32_t nold = frame == 0 ? 0u : ac.sample_n[pix];
  if (frame == 0) ac.sample_n[pix] = 0u;
  uint32_t nn = nold + n;
  if (!(nn > nold) || isinf3(acc)) return;
  float ri = 1.f / (float)nn, oi = (float)nold * ri;
  float* r = ac.result + (size_t)pix * 3;
  r[0] = fmaxf(ri * acc.x + oi * r[0], 0.f);
  r[1] = fmaxf(ri * acc.y + oi * r[1], 0.f);
  r[2] = fmaxf(ri * acc.z + oi * r[2], 0.f);
  ac.sample_n[pix] = nn;
  const float nf = (float)n;
  float4 g[3] = {ac.rec.gpos[sid], ac.rec.gnrm[sid], ac.rec.gdif[sid]};
  float* dst[3] = {ac.gpos + (size_t)pix * 3, ac.gnrm + (size_t)pix * 3, ac.gdif + (size_t)pix * 3};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    f3 w = ld3(g[k]) * nf;
    dst[k][0] = ri * w.x + oi * dst[k][0];
    dst[k][1] = ri * w.y + oi * dst[k][1];
    dst[k][2] = ri * w.z + oi * dst[k][2];
  }
  if (ac.sq) {                                     // squareResult (AdapEmp.cu:927-944)
    float* q = ac.sq + (size_t)pix * 3;
    q[0] = fmaxf(ri * (acc.x * acc.x) + oi * q[0], 0.f);
    q[1] = fmaxf(ri * (acc.y * acc.y) + oi * q[1], 0.f);
    q[2] = fmaxf(ri * (acc.z * acc.z) + oi * q[2], 0.f);
  }
}

// Append with one atomic per workgroup of B threads, grouped by a 3-bit key
// (the ray direction's octant): within the workgroup's share of each queue the
// lanes come key by key, so that a wave of the next traversal holds rays of few
// octants.  k0 / k1: the key of a lane appending to queue 0 / 1.
template <int B>
__device__ __forceinline__ void block_append2_oct(bool p0, uint32_t k0, bool p1, uint32_t k1, uint32_t* ctr0,
                                                  uint32_t* ctr1, uint32_t& s0, uint32_t& s1) {
  constexpr int W = B / 64;
  __shared__ uint32_t c0[8][W], c1[8][W], base0, base1;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t r0 = 0, r1 = 0;
#pragma unroll
  for (uint32_t b = 0; b < 8; b++) {
    const unsigned long long m0 = __ballot(p0 && k0 == b), m1 = __ballot(p1 && k1 == b);
    if (lane == 0) { c0[b][wid] = (uint32_t)__popcll(m0); c1[b][wid] = (uint32_t)__popcll(m1); }
    if (k0 == b) r0 = (uint32_t)__popcll(m0 & below);
    if (k1 == b) r1 = (uint32_t)__popcll(m1 & below);
  }
  __syncthreads();
  if (threadIdx.x == 0) {                  // exclusive prefix over (key, wave), then one atomic per queue
    uint32_t t0 = 0, t1 = 0;
    for (int b = 0; b < 8; b++)
      for (int w = 0; w < W; w++) {
        const uint32_t a = c0[b][w], c = c1[b][w];
        c0[b][w] = t0; c1[b][w] = t1;
        t0 += a; t1 += c;
      }
    base0 = t0 ? atomicAdd(ctr0, t0) : 0u;
    base1 = t1 ? atomicAdd(ctr1, t1) : 0u;
  }
  __syncthreads();
  s0 = p0 ? base0 + c0[k0 & 7u][wid] + r0 : 0u;
  s1 = p1 ? base1 + c1[k1 & 7u][wid] + r1 : 0u;
}
__device__ __forceinline__ uint32_t octant(float4 d) {
  return (d.x < 0.f ? 1u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 4u : 0u);
}

// One eye-path vertex (pt_traceSample_kernel, src/cudaRTBDPTStream.cu:365-543,
// and the connection set-up of pt_connectEyeLightPath_kernel, :665-723):
// emission, BRDF sample, Russian roulette, termination into the accumulators,
// and for a surviving diffuse bounce the light-vertex connection's shadow ray.
// p holds the path before the bounce and, when `survive`, after it.
__device__ __forceinline__ void shade_path(const DevScene& sc, const FrameArgs& fa, const FrameCounters* fc,
                                           const LightVert* __restrict__ lv, const Accum& ac, PathState& p, float4 h,
                                           bool& survive, bool& emit, ShadowRay& sr) {
  const uint32_t pix = p.pix, sid = p.sid;
  uint32_t depth = p.depth;
  const uint32_t frame = frame_base(fa) + p.fk;
  const uint32_t slot = (frame / 3u) % (uint32_t)kLightSlots;
  const uint32_t lvN = fc->light_n[slot];
  const LightVert* __restrict__ lvb = lv + (size_t)slot * fa.lv_stride;
  uint64_t st = p.rng;
  const float4 ro = p.ro, rd = p.rd, m4 = p.mul, s4 = p.smp, a4 = p.acc;
  float4 o_out, d_out, mul_out, smp_out, acc_out;
  float4& so = sr.o;
  float4& sd = sr.d;
  float4& sc4 = sr.c;
  const uint64_t inc = ((uint64_t)local_pixel(fa, pix) << 1) | 1u;
  f3 orig = ld3(ro), dir = ld3(rd), mul = ld3(m4), psample = ld3(s4), acc = ld3(a4);
  uint32_t N = __float_as_uint(a4.w);
  uint32_t leaf = __float_as_uint(h.y);
  bool term = false;
  if (leaf == kInvalid) {
    // an escaping eye path gathers the environment's radiance through its throughput
    if (sc.env_w) psample = psample + vmul(env_radiance(sc, dir), mul);
    acc = acc + psample; N++; term = true;
  } else {
    float w = (1.f - h.z) - h.w;
    const bool gbuf = (ac.variant & kVariantGbuf) && depth == 0;
    Surf s = fetch_surf(sc, leaf, w, h.z, h.w, gbuf);
    f3 nl = dot(s.norm, dir) < 0.f ? s.norm : s.norm * -1.f;
    if (gbuf) {                                          // AdapBilatAtomic.cu:650-655
      ac.rec.gpos[sid] = make_float4(s.pos.x, s.pos.y, s.pos.z, 0.f);
      ac.rec.gnrm[sid] = make_float4(nl.x, nl.y, nl.z, 0.f);
      ac.rec.gdif[sid] = make_float4(s.diff.x, s.diff.y, s.diff.z, 0.f);
    }
    const bool pt = (ac.variant & kVariantPT) != 0;
    Bounce b = sample_bounce(st, inc, orig, dir, h.x, s, nl, RT_EYE, (ac.variant & kVariantReflOffset) != 0, pt);
    psample = psample + vmul(s.emissive, mul);
    f3 inMul = mul, inDir = dir * -1.f;
    float contrib = length(mul) * length(b.mul);
    if (length(s.emissive) > 0.f) contrib = 0.f;
    // the regenerating path tracer also ends a diffuse-to-specular step
    // (cudaRTPTStreamRegen.cu:241)
    if ((ac.variant & kVariantPTRegen) && p.type == RT_DIFF && b.type == RT_SPEC) contrib = 0.f;
    const uint32_t lim = p.lim ? p.lim : fa.eye_depth;
    if (uniform01(st, inc) > contrib || depth + 1 >= lim) {
      acc = acc + psample; N++; term = true;
    } else {
      mul = vmul(mul, b.mul);
      dir = b.dir;
      depth++;
      orig = b.orig;
      p.type = b.type;
      if (!pt && b.type == RT_DIFF && lvN > 0) {
        f3 pos = orig + nl * (float)kBiasEps;
        uint32_t li = (uint32_t)(uniform01(st, inc) * (float)lvN);
        if (li >= lvN) li = lvN - 1;
        const LightVert& L = lvb[li];
        f3 lpos = mk3(L.pos[0], L.pos[1], L.pos[2]);
        f3 lirr = mk3(L.irrad[0], L.irrad[1], L.irrad[2]);
        f3 toL = normalize(lpos - pos);
        float dist = length(lpos - pos);
        if (length(lirr) > 0.f && dot(nl, toL) > 0.f) {
          f3 irrad, irradDir;
          if ((double)dist > kFltEps) {
            irrad = shading(toL * -1.f, mk3(L.irradDir[0], L.irradDir[1], L.irradDir[2]), lirr,
                            mk3(L.norm[0], L.norm[1], L.norm[2]), mk3(L.diff[0], L.diff[1], L.diff[2]), L.metallic,
                            L.roughness, L.specular, 1.f, 1.f) +
                    mk3(L.emissive[0], L.emissive[1], L.emissive[2]);
            irradDir = toL;
          } else {
            irrad = lirr;
            irradDir = mk3(L.irradDir[0], L.irradDir[1], L.irradDir[2]) * -1.f;
          }
          f3 c = vmax0(shading(inDir, irradDir, irrad, nl, s.diff, s.metal, s.rough, s.spec, 1.f - s.trans, 1.f));
          if (length(c) > 0.f) {
            f3 pay = vmul(c, inMul);
            emit = true;
            so = make_float4(pos.x, pos.y, pos.z, (float)shadow_max((double)dist, fa.flags));
            sd = make_float4(toL.x, toL.y, toL.z, (float)kBiasEps);
            sc4 = make_float4(pay.x, pay.y, pay.z, 0.f);
          }
        }
      }
    }
  }
  if (term) {
    if (ac.term_bounce) {                                 // regeneration: where and in which state it ended
      ac.term_bounce[sid] = depth;
      ac.term_rng[sid] = st;
    }
    if (ac.apply == 0) apply_pixel(ac.result, ac.sample_n, pix, frame, acc, N);
    else if (ac.apply == 1) apply_pixel_gbuf(ac, pix, frame, acc, N, sid);
    else ac.rec.acc[sid] = make_float4(acc.x, acc.y, acc.z, __uint_as_float(N));
  } else {
    survive = true;
    o_out = make_float4(orig.x, orig.y, orig.z, kInfF);
    d_out = make_float4(dir.x, dir.y, dir.z, (float)kBiasEps);
    mul_out = make_float4(mul.x, mul.y, mul.z, 0.f);
    smp_out = make_float4(psample.x, psample.y, psample.z, 0.f);
    acc_out = make_float4(acc.x, acc.y, acc.z, __uint_as_float(N));
  }
  if (survive) {
    p.ro = o_out;
    p.rd = d_out;
    p.mul = mul_out;
    p.smp = smp_out;
    p.acc = acc_out;
    p.depth = depth;
  }
  p.rng = st;
}

// ---- Disney-BRDF path tracer (extension; no reference counterpart) ----------
// The principled BRDF of Burley (SIGGRAPH 2012 course notes and brdf.txt),
// isotropic, specular tint 0, opaque, with next-event estimation to the
// emissive triangles and power-heuristic MIS against the BRDF samples.  The
// same operations, in the same order and drawing the same uniforms, as the
// oracle's FP64 restatement (disney_path).
struct DisneyMat {
  f3 base;
  float metallic, subsurface, specular, roughness, sheen, sheenTint, clearcoat, clearcoatGloss;
};
__device__ __forceinline__ float schlick_w(float c) {
  float m = fminf(fmaxf(1.f - c, 0.f), 1.f);
  return (m * m) * (m * m) * m;
}
__device__ __forceinline__ float gtr1(float nh, float a) {
  if (a >= 1.f) return (float)(1.0 / kPiD);
  const float a2 = a * a, t = 1.f + (a2 - 1.f) * nh * nh;
  return (a2 - 1.f) / (kPiF * logf(a2) * t);
}
__device__ __forceinline__ float gtr2(float nh, float a) {
  const float a2 = a * a, t = 1.f + (a2 - 1.f) * nh * nh;
  return a2 / (kPiF * t * t);
}
__device__ __forceinline__ float smith_g(float n, float a) {
  const float a2 = a * a, b = n * n;
  return 1.f / (n + sqrtf(a2 + b - a2 * b));
}
__device__ __forceinline__ float disney_alpha(const DisneyMat& m) { return fmaxf(m.roughness * m.roughness, 1e-3f); }
__device__ __forceinline__ float disney_cc_alpha(const DisneyMat& m) {
  return 0.1f + (0.001f - 0.1f) * m.clearcoatGloss;
}
__device__ __forceinline__ void disney_lobes(const DisneyMat& m, float& pd, float& ps, float& pc) {
  const float wd = 1.f - m.metallic, ws = 1.f, wc = 0.25f * m.clearcoat, t = (wd + ws) + wc;
  pd = wd / t; ps = ws / t; pc = wc / t;
}
__device__ f3 disney_f(const DisneyMat& m, f3 n, f3 wo, f3 wi) {
  const float nl = dot(n, wi), nv = dot(n, wo);
  if (nl <= 0.f || nv <= 0.f) return mk3(0.f, 0.f, 0.f);
  const f3 hv = wi + wo;
  const float hl = length(hv);
  if (!(hl > 0.f)) return mk3(0.f, 0.f, 0.f);
  const f3 h = hv / hl;
  const float nh = dot(n, h), lh = dot(wi, h);
  const float lum = 0.3f * m.base.x + 0.6f * m.base.y + 0.1f * m.base.z;
  const f3 tint = lum > 0.f ? m.base / lum : mk3(1.f, 1.f, 1.f);
  const f3 one = mk3(1.f, 1.f, 1.f);
  const f3 cspec0 = one * (0.08f * m.specular) + (m.base - one * (0.08f * m.specular)) * m.metallic;
  const f3 csheen = one + (tint - one) * m.sheenTint;
  const float fl = schlick_w(nl), fv = schlick_w(nv), fh = schlick_w(lh);
  const float fd90 = 0.5f + 2.f * lh * lh * m.roughness;
  const float fd = (1.f + (fd90 - 1.f) * fl) * (1.f + (fd90 - 1.f) * fv);
  const float fss90 = lh * lh * m.roughness;
  const float fss = (1.f + (fss90 - 1.f) * fl) * (1.f + (fss90 - 1.f) * fv);
  const float ss = 1.25f * (fss * (1.f / (nl + nv) - 0.5f) + 0.5f);
  const float a = disney_alpha(m);
  const float ds = gtr2(nh, a);
  const f3 fs = cspec0 + (one - cspec0) * fh;
  const float gs = smith_g(nl, a) * smith_g(nv, a);
  const f3 fsheen = csheen * (fh * m.sheen);
  const float dr = gtr1(nh, disney_cc_alpha(m));
  const float fr = 0.04f + (1.f - 0.04f) * fh;
  const float gr = smith_g(nl, 0.25f) * smith_g(nv, 0.25f);
  const f3 diffuse = (m.base * ((float)(1.0 / kPiD) * (fd + (ss - fd) * m.subsurface)) + fsheen) * (1.f - m.metallic);
  return (diffuse + fs * (gs * ds)) + one * (0.25f * m.clearcoat * gr * fr * dr);
}
__device__ float disney_pdf(const DisneyMat& m, f3 n, f3 wo, f3 wi) {
  const float nl = dot(n, wi), nv = dot(n, wo);
  if (nl <= 0.f || nv <= 0.f) return 0.f;
  const f3 hv = wi + wo;
  const float hl = length(hv);
  if (!(hl > 0.f)) return 0.f;
  const f3 h = hv / hl;
  const float nh = dot(n, h), vh = dot(wo, h);
  if (vh <= 0.f) return 0.f;
  float pd, ps, pc;
  disney_lobes(m, pd, ps, pc);
  float pdf = pd * nl * (float)(1.0 / kPiD);
  pdf = pdf + ps * gtr2(nh, disney_alpha(m)) * nh / (4.f * vh);
  pdf = pdf + pc * gtr1(nh, disney_cc_alpha(m)) * nh / (4.f * vh);
  return pdf;
}
__device__ bool disney_sample(const DisneyMat& m, f3 n, f3 wo, float u0, float u1, float u2, f3& wi) {
  float pd, ps, pc;
  disney_lobes(m, pd, ps, pc);
  f3 u, v;
  basis(n, u, v);
  const float phi = (float)(2.0 * kPiD * (double)u1);
  if (u0 < pd) {
    const float ct = sqrtf(u2), st = sqrtf(1.f - u2);
    float sp, cp;
    sincos_def(phi, sp, cp);
    wi = normalize((n * ct + (u * st) * cp) + (v * st) * sp);
  } else {
    const bool spec = u0 < pd + ps;
    const float a = spec ? disney_alpha(m) : disney_cc_alpha(m);
    float ct;
    if (spec) ct = sqrtf((1.f - u2) / (1.f + (a * a - 1.f) * u2));
    else ct = a >= 1.f ? sqrtf(1.f - u2) : sqrtf((1.f - powf(a * a, 1.f - u2)) / (1.f - a * a));
    const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
    float sp, cp;
    sincos_def(phi, sp, cp);
    const f3 h = (n * ct + (u * st) * cp) + (v * st) * sp;
    wi = normalize(h * (2.f * dot(wo, h)) - wo);
  }
  return dot(n, wi) > 0.f;
}
__device__ __forceinline__ DisneyMat disney_of(const DevScene& sc, uint32_t leaf, const Surf& s) {
  DisneyMat m;
  m.base = s.diff; m.metallic = s.metal; m.specular = s.spec; m.roughness = s.rough;
  const uint32_t mi = __float_as_uint(sc.surf[8 * (size_t)leaf + 7].w);
  const DevMaterial& d = sc.mats[mi];
  m.subsurface = d.subsurface; m.sheen = d.sheen; m.sheenTint = d.sheenTint;
  m.clearcoat = d.clearcoat; m.clearcoatGloss = d.clearcoatGloss;
  return m;
}
// Solid-angle density of the light sampler at the point at distance t along
// unit d on leaf triangle `leaf`: 1 / n_lights, uniform by area, two faces.
__device__ __forceinline__ float light_pdf(const DevScene& sc, uint32_t leaf, float t, f3 d) {
  const f3 ng = cross(ld3(sc.tris[3 * (size_t)leaf + 1]), ld3(sc.tris[3 * (size_t)leaf + 2]));
  const float a2 = length(ng);
  const float c = fabsf(dot(ng, d)) / a2;
  if (!(c > 0.f) || !(a2 > 0.f)) return 0.f;
  return t * t / (c * (a2 * 0.5f) * (float)sc.n_light_tris);
}
// One bounce of the Disney path tracer: emission (MIS-weighted against the
// previous bounce's light sample), the depth bound, Russian roulette, the light
// sample (a shadow ray adding to the path's accumulator without a count) and
// the BRDF sample.  smp.w carries the BRDF density of the ray being traced.
__device__ __forceinline__ void shade_disney(const DevScene& sc, const FrameArgs& fa, const Accum& ac, PathState& p,
                                             float4 h, bool& survive, bool& emit, ShadowRay& sr) {
  const uint32_t pix = p.pix, sid = p.sid;
  uint32_t depth = p.depth;
  const uint32_t frame = frame_base(fa) + p.fk;
  uint64_t st = p.rng;
  const uint64_t inc = ((uint64_t)local_pixel(fa, pix) << 1) | 1u;
  f3 orig = ld3(p.ro), dir = ld3(p.rd), T = ld3(p.mul), L = ld3(p.smp), acc = ld3(p.acc);
  const float prev_pdf = p.smp.w;
  uint32_t N = __float_as_uint(p.acc.w);
  const uint32_t leaf = __float_as_uint(h.y);
  bool term = false, dead = false;
  float next_pdf = 0.f;
  if (leaf == kInvalid) {
    if (sc.env_w) L = L + vmul(T, env_radiance(sc, dir));
    term = true;
  } else {
    const float w = (1.f - h.z) - h.w;
    const Surf s = fetch_surf(sc, leaf, w, h.z, h.w, false);
    if (length(s.emissive) > 0.f) {
      float wt = 1.f;
      if (depth > 0) {
        const float pl = light_pdf(sc, leaf, h.x, dir);
        wt = prev_pdf * prev_pdf / (prev_pdf * prev_pdf + pl * pl);
      }
      L = L + vmul(T, s.emissive) * wt;
    }
    const uint32_t lim = p.lim ? p.lim : fa.eye_depth;
    if (depth + 1 >= lim) term = true;
    if (!term && depth >= 3) {
      const float q = fminf(0.95f, fmaxf(T.x, fmaxf(T.y, T.z)));
      if (uniform01(st, inc) >= q) term = true;
      else T = T / q;
    }
    if (!term) {
      const f3 nl = dot(s.norm, dir) < 0.f ? s.norm : s.norm * -1.f;
      const f3 pos = orig + dir * h.x, wo = dir * -1.f;
      const DisneyMat m = disney_of(sc, leaf, s);
      const uint32_t nL = sc.n_light_tris;
      if (nL > 0) {
        const float l0 = uniform01(st, inc);
        float l1 = uniform01(st, inc), l2 = uniform01(st, inc);
        uint32_t li = (uint32_t)(l0 * (float)nL);
        if (li >= nL) li = nL - 1;
        if (l1 + l2 > 1.f) { l1 = 1.f - l1; l2 = 1.f - l2; }
        const uint32_t lt = sc.light_tris[li];
        const float4 q0 = sc.tris[3 * (size_t)lt], q1 = sc.tris[3 * (size_t)lt + 1], q2 = sc.tris[3 * (size_t)lt + 2];
        const f3 lp = (ld3(q0) + ld3(q1) * l1) + ld3(q2) * l2;
        const f3 so = pos + nl * (float)kBiasEps;
        const f3 toL = lp - so;
        const float dist = length(toL);
        if ((double)dist > kFltEps) {
          const f3 wi = toL / dist;
          const float cs = dot(nl, wi);
          const float pl = light_pdf(sc, lt, dist, wi);
          if (cs > 0.f && pl > 0.f) {
            const f3 f = disney_f(m, nl, wo, wi);
            const float pb = disney_pdf(m, nl, wo, wi);
            const float wt = pl * pl / (pl * pl + pb * pb);
            const f3 le = ld3(sc.surf[8 * (size_t)lt + 4]);
            const f3 c = vmul(vmul(T, f), le) * (cs * wt / pl);
            if (length(c) > 0.f) {
              emit = true;
              sr.o = make_float4(so.x, so.y, so.z, (float)((double)dist - 2.0 * kBiasEps));
              sr.d = make_float4(wi.x, wi.y, wi.z, (float)kBiasEps);
              sr.c = make_float4(c.x, c.y, c.z, 0.f);
            }
          }
        }
      }
      const float b0 = uniform01(st, inc), b1 = uniform01(st, inc), b2 = uniform01(st, inc);
      f3 wi;
      const bool ok = disney_sample(m, nl, wo, b0, b1, b2, wi);
      const float pdf = ok ? disney_pdf(m, nl, wo, wi) : 0.f;
      if (!(pdf > 0.f)) {
        // no continuation; a pending light sample still needs the slot: the
        // path lives one more, empty segment (maximum distance 0, throughput 0)
        if (emit) dead = true;
        else term = true;
      } else {
        T = vmul(T, disney_f(m, nl, wo, wi)) * (dot(nl, wi) / pdf);
        next_pdf = pdf;
        orig = pos + wi * (float)kBiasEps;
        dir = wi;
      }
    }
  }
  if (term) {
    acc = acc + L;
    N++;
    if (ac.apply == 0) apply_pixel(ac.result, ac.sample_n, pix, frame, acc, N);
    else ac.rec.acc[sid] = make_float4(acc.x, acc.y, acc.z, __uint_as_float(N));
  } else {
    survive = true;
    if (dead) T = mk3(0.f, 0.f, 0.f);
    p.ro = make_float4(orig.x, orig.y, orig.z, dead ? 0.f : kInfF);
    p.rd = make_float4(dir.x, dir.y, dir.z, (float)kBiasEps);
    p.mul = make_float4(T.x, T.y, T.z, 0.f);
    p.smp = make_float4(L.x, L.y, L.z, next_pdf);
    p.acc = make_float4(acc.x, acc.y, acc.z, __uint_as_float(N));
    p.depth = depth + 1;
    p.type = RT_DIFF;
  }
  p.rng = st;
}

// The shading pass of one bounce over the live paths: survivors compacted into
// nxt, connection shadow rays into sb (contribution .w = the survivor's slot).
// Workgroup of the shading pass: its compaction takes one atomic per counter
// per workgroup.
constexpr int kShadeBlock = 256;
// Five waves per SIMD: 96 VGPRs without spills; left to itself the compiler
// takes 97 and four waves, and the pass is then 5-7 % slower.
// DISNEY: the Disney path tracer's bounce (shade_disney) in an instance of its
// own, so that its larger live set does not spill the BDPT instance's.
template <bool DISNEY>
__global__ __launch_bounds__(kShadeBlock, DISNEY ? 4 : 5) void k_shade(DevScene sc, FrameArgs fa, PathBuf cur,
                                                                       PathBuf nxt,
                                                  const float4* __restrict__ hits, FrameCounters* __restrict__ fc,
                                                  int bounce, ShadowBuf sb, const LightVert* __restrict__ lv,
                                                  Accum ac, int eye) {
  const uint32_t i = blockIdx.x * kShadeBlock + threadIdx.x;
  const uint32_t count = fc->active[bounce];
  if (blockIdx.x * kShadeBlock >= count) return;               // whole workgroup idle
  bool survive = false, emit = false;
  PathState p;
  ShadowRay sr;
  sr.o = sr.d = sr.c = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i < count) {
    uint32_t dk;
    if (eye) {
      p = eye_slot(fa, i, dk);
    } else {
      p.ro = cur.ray_o[i];
      p.rd = cur.ray_d[i];
      p.pix = cur.pix[i];
      p.sid = cur.sid[i];
      dk = cur.depth[i];
      p.rng = cur.rng[i];
    }
    unpack_dk(p, dk);
    if (dk & kFreshPath) {
      p.mul = make_float4(1.f, 1.f, 1.f, 0.f);
      p.smp = make_float4(0.f, 0.f, 0.f, 0.f);
      p.acc = p.smp;
    } else {
      p.mul = cur.mul[i];
      p.smp = cur.smp[i];
      p.acc = cur.acc[i];
    }
    if constexpr (DISNEY) shade_disney(sc, fa, ac, p, hits[i], survive, emit, sr);
    else shade_path(sc, fa, fc, lv, ac, p, hits[i], survive, emit, sr);
  }
  uint32_t slot = 0, sslot = 0;
  block_append2_oct<kShadeBlock>(survive, survive ? octant(p.rd) : 8u, emit, emit ? octant(sr.d) : 8u,
                                 &fc->active[bounce + 1], &fc->shadow[bounce], slot, sslot);
  if (survive) {
    nxt.ray_o[slot] = p.ro;
    nxt.ray_d[slot] = p.rd;
    nxt.pix[slot] = p.pix;
    nxt.sid[slot] = p.sid;
    nxt.depth[slot] = pack_dk(p);
    nxt.mul[slot] = p.mul;
    nxt.smp[slot] = p.smp;
    nxt.acc[slot] = p.acc;
    nxt.rng[slot] = p.rng;
  }
  if (emit) {
    sr.c.w = __uint_as_float(slot | (DISNEY ? kShadowNoCount : 0u));
    sb.ray_o[sslot] = sr.o;
    sb.ray_d[sslot] = sr.d;
    sb.contrib[sslot] = sr.c;
  }
}
hipError_t launch_shade(const DevScene& sc, const FrameArgs& fa, PathBuf cur, PathBuf nxt, const float4* hits,
                        FrameCounters* fc, int bounce, uint32_t max_n, ShadowBuf sb, const LightVert* lv,
                        const Accum& ac, hipStream_t st, int eye) {
  if (max_n == 0) return hipSuccess;
  if (!pathbuf_ok(cur) || !pathbuf_ok(nxt) || !hits || !fc || !lv || !sb.ray_o || !sb.ray_d || !sb.contrib ||
      !ac.result || !ac.sample_n)
    return hipErrorInvalidValue;
  if (ac.apply == 1 && (!rec_ok(ac.rec) || !ac.gpos || !ac.gnrm || !ac.gdif)) return hipErrorInvalidValue;
  if (ac.apply == 2 && !ac.rec.acc) return hipErrorInvalidValue;
  if ((ac.variant & kVariantGbuf) && !rec_ok(ac.rec)) return hipErrorInvalidValue;
  const uint32_t grid = (max_n + kShadeBlock - 1) / kShadeBlock;
  if (ac.variant & kVariantDisney)
    k_shade<true><<<grid, kShadeBlock, 0, st>>>(sc, fa, cur, nxt, hits, fc, bounce, sb, lv, ac, eye);
  else
    k_shade<false><<<grid, kShadeBlock, 0, st>>>(sc, fa, cur, nxt, hits, fc, bounce, sb, lv, ac, eye);
  return hipGetLastError();
}

template <int STK, int MINW = 0, bool Q = false>
__global__ __launch_bounds__(kBlock, MINW) void k_shadow(DevScene sc, ShadowBuf sb, PathBuf nxt,
                                                   const FrameCounters* __restrict__ fc, int bounce,
                                                   DevStats* __restrict__ stats) {
  __shared__ uint64_t lstk[STK * kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t count = fc->shadow[bounce];
  uint32_t nn = 0, nt = 0;
  if (i < count) {
    float4 a = sb.ray_o[i], b = sb.ray_d[i], c = sb.contrib[i];
    TraceHit h;
    if (!traverse4<true, STK, Q>(sc, ld3(a), ld3(b), a.w, b.w, false, lds_column(lstk), h, nn, nt)) {
      const uint32_t bits = __float_as_uint(c.w), slot = bits & ~kShadowNoCount;
      float4 acc = nxt.acc[slot];
      acc.x = acc.x + c.x;
      acc.y = acc.y + c.y;
      acc.z = acc.z + c.z;
      acc.w = __uint_as_float(__float_as_uint(acc.w) + ((bits & kShadowNoCount) ? 0u : 4u));
      nxt.acc[slot] = acc;
    }
  }
  if (stats) {
    uint32_t sn = wave_sum(nn), st = wave_sum(nt);
    if ((threadIdx.x & 63) == 0 && (sn | st)) {
      atomicAdd(&stats->sh_nodes, (unsigned long long)sn);
      atomicAdd(&stats->sh_tris, (unsigned long long)st);
    }
  }
}
hipError_t launch_shadow(const DevScene& sc, ShadowBuf sb, PathBuf nxt, const FrameCounters* fc, int bounce,
                         uint32_t max_n, int lds_stack, int quant, DevStats* stats, hipStream_t st) {
  if (max_n == 0) return hipSuccess;
  if (!pathbuf_ok(nxt) || !sb.ray_o || !sb.ray_d || !sb.contrib || !fc || !sc.tris ||
      !(quant ? (const void*)sc.nodes4q : (const void*)sc.nodes4))
    return hipErrorInvalidValue;
  dim3 grid((max_n + kBlock - 1) / kBlock), block(kBlock);
#define CPT_SH(S, W, Q) k_shadow<S, W, Q><<<grid, block, 0, st>>>(sc, sb, nxt, fc, bounce, stats)
  if (quant) {
#define CPT_SH_Q(S, W) CPT_SH(S, W, true)
#define CPT_SH_F(S, W) CPT_SH(S, W, false)
    CPT_VARIANT_SWITCH(lds_stack, CPT_SH_Q)
  } else {
    CPT_VARIANT_SWITCH(lds_stack, CPT_SH_F)
  }
#undef CPT_SH_Q
#undef CPT_SH_F
#undef CPT_SH
  return hipGetLastError();
}

// ---------------------------------------------------------------- path regeneration
// cudaRTPTStreamRegen (src/cudaRTPTStreamRegen.cu:300-478): a slot whose path
// ends while other paths of the frame are still being traced restarts once,
// for the same pixel with its RNG state continuing (pt_reinitRegenPathStream_kernel),
// and the frame stops when its last first-generation path ends; a restarted
// path still running then adds what it gathered (pt_applyPathQueueResult_kernel).
// While the stream holds the whole frame (W*H <= 640,000 there), a path that
// ended at bounce b runs bounces b+1 .. T of the restart, T the frame's last
// first-generation bounce, which no other path influences.  So the restarts
// are traced as a second wavefront after the first: tend[k] = T of frame k,
// and the restart's depth bound is min(eye depth, T - b).

// tend[k] = max over the frame's slots of the bounce its path ended at.
__global__ __launch_bounds__(kBlock) void k_term_max(const uint32_t* __restrict__ tb, uint32_t wh,
                                                     uint32_t* __restrict__ tend) {
  __shared__ uint32_t wmax[kBlock / 64];
  const uint32_t k = blockIdx.y;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t v = i < wh ? tb[(size_t)k * wh + i] : 0u;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (int w = 0; w < kBlock / 64; w++) m = max(m, wmax[w]);
    if (m) atomicMax(&tend[k], m);
  }
}

// The restarts: sample slot sid of frame k when its path ended before tend[k],
// visited pixel-major as k_gen_eye lays out first paths; workgroups of 1024 so
// that the append takes a quarter of the atomics on the one counter (the
// restart pass took 3.9 ms at 1280 x 720 x 64 frames with 256).
constexpr int kRegenBlock = 1024;
__global__ __launch_bounds__(kRegenBlock) void k_gen_regen(FrameArgs fa, PathBuf pb, const uint32_t* __restrict__ tb,
                                                      const uint64_t* __restrict__ trng,
                                                      const uint32_t* __restrict__ tend,
                                                      const float4* __restrict__ rec_acc,
                                                      FrameCounters* __restrict__ fc) {
  __shared__ uint32_t wc[kRegenBlock / 64], base;
  const uint32_t wh = fa.width * fa.height;
  const uint32_t i = blockIdx.x * kRegenBlock + threadIdx.x;
  bool take = false;
  PathState p;
  if (i < wh * fa.nframes) {
    const uint32_t K = fa.nframes, pix = i / K, k = i - pix * K, sid = k * wh + pix;
    const uint32_t b = tb[sid], T = tend[k];
    if (b < T) {
      take = true;
      p = eye_path_at(fa, pix, sid, trng[sid]);
      p.fk = k;
      p.lim = min(T - b, fa.eye_depth);
      p.acc = rec_acc[sid];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long m = __ballot(take);
  if (lane == 0) wc[wid] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kRegenBlock / 64; w++) t += wc[w];
    base = t ? atomicAdd(&fc->active[0], t) : 0u;
  }
  __syncthreads();
  if (take) {
    uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < wid; w++) slot += wc[w];
    pb.ray_o[slot] = p.ro;
    pb.ray_d[slot] = p.rd;
    pb.pix[slot] = p.pix;
    pb.sid[slot] = p.sid;
    pb.depth[slot] = pack_dk(p);
    pb.mul[slot] = p.mul;
    pb.smp[slot] = p.smp;
    pb.acc[slot] = p.acc;
    pb.rng[slot] = p.rng;
  }
}
hipError_t launch_regen(const FrameArgs& fa, PathBuf pb, const Accum& ac, FrameCounters* fc, hipStream_t st) {
  const uint32_t wh = fa.width * fa.height;
  const uint64_t n = (uint64_t)wh * fa.nframes;
  if (n == 0) return hipSuccess;
  if (!pathbuf_ok(pb) || !fc || !ac.term_bounce || !ac.term_rng || !ac.tend || !ac.rec.acc || fa.nframes > 255 ||
      fa.views)
    return hipErrorInvalidValue;
  k_fill_n_u32<<<1, kBlock, 0, st>>>(ac.tend, fa.nframes, 0u);
  k_term_max<<<dim3((wh + kBlock - 1) / kBlock, fa.nframes), kBlock, 0, st>>>(ac.term_bounce, wh, ac.tend);
  k_frame_begin<<<1, 64, 0, st>>>(fc, 0u);
  k_gen_regen<<<(uint32_t)((n + kRegenBlock - 1) / kRegenBlock), kRegenBlock, 0, st>>>(fa, pb, ac.term_bounce, ac.term_rng, ac.tend,
                                                                         ac.rec.acc, fc);
  return hipGetLastError();
}

// ---------------------------------------------------------------- "Path Regen" BDPT
// The reference's megakernel BDPT renderer (src/cudaRTBDPTRegen.cu, the
// "Path Regen" baseline), one thread per pixel.
//  - k_regen_light, every third frame: 128 light paths of 5 vertices each fill
//    a 640-vertex cache (bdpt_light_kernel :314-330, bdpt_light_ray
//    :80-309).  A path starts on a random emissive triangle (vertex: irrad =
//    emissive, no direction) and then carries the throughput (1, 1, 1), not
//    the emission (:179, :307); later vertices store the throughput they
//    arrived with and the hit's interpolated normal; a miss stores an empty
//    vertex, and the next vertex starts a new path, as does a throughput
//    below 1e-4.
//  - k_regen_eye: pt0_normalRay's recursion (:433-625) as a loop over at most
//    eye_depth levels with the per-level factors kept in registers: the BRDF
//    bounce (refraction past the critical angle reflects, :501-505), Russian
//    roulette on the pixel weight, whose failure restarts a camera path for
//    the same pixel (:562-571) and whose success connects a diffuse bounce to a
//    uniformly chosen cache vertex; then the radiance summed back up the
//    levels, r = f * r' + emission + connection, banked where a camera path
//    starts (:598-606), and the pixel's running mean over the count of camera
//    paths plus 4 per contributing connection (:655-664).
constexpr uint32_t kRegenLightN = kRegenLightVerts, kRegenLightDepth = 5, kRegenMaxLevels = 8;
constexpr uint32_t kDomainRegenLight = 4u, kDomainRegenEye = 5u;

__global__ __launch_bounds__(kRegenLightN / kRegenLightDepth) void k_regen_light(DevScene sc, FrameArgs fa,
                                                                                  LightVert* __restrict__ lv) {
  __shared__ uint64_t lstk_all[CPT_LDS_STACK * kBlock];   // column layout: the stride is kBlock
  LdsU64* lstk = lds_column(lstk_all);
  const uint32_t x = threadIdx.x;
  const uint64_t inc = ((uint64_t)x << 1) | 1u;
  uint64_t st;
  pcg32_seed(st, inc, stream_key(fa.seed, fa.frame0, kDomainRegenLight, 0u));
  bool valid = false;
  f3 o = mk3(0.f, 0.f, 0.f), d = o, irr = o;
  const Surf none = {o, o, o, o, 0.f, 0.f, 0.f, 0.f};
  for (uint32_t k = 0; k < kRegenLightDepth; k++) {
    f3 mulT = mk3(1.f, 1.f, 1.f), no = o, nd = d;
    bool nvalid = false;
    LightVert& V = lv[x * kRegenLightDepth + k];
    store_lv(V, o * 0.f, o * 0.f, o * 0.f, o * 0.f, none);
    if (!valid || length(irr) < (float)kBiasEps) {
      if (sc.n_light_tris > 0) {                         // a new path on an emitter (:94-158)
        uint32_t id = (uint32_t)(uniform01(st, inc) * (float)sc.n_light_tris);
        if (id >= sc.n_light_tris) id = sc.n_light_tris - 1;
        float lw = uniform01(st, inc), lu = uniform01(st, inc);
        if (lw + lu > 1.f) { lw = 1.f - lw; lu = 1.f - lu; }
        const float lvv = (1.f - lw) - lu;
        const Surf s = fetch_surf(sc, sc.light_tris[id], lw, lu, lvv, true);
        f3 u, v;
        basis(s.norm, u, v);
        const float r1 = (float)(2.0 * kPiD * (double)uniform01(st, inc));
        const float r2cos = sqrtf(uniform01(st, inc));
        const float r2sin = 1.f - r2cos * r2cos;
        float s1, c1;
        sincos_def(r1, s1, c1);
        nd = normalize((s.norm * r2cos + (u * r2sin) * c1) + (v * r2sin) * s1);
        no = s.pos + s.norm * (float)kBiasEps;
        nvalid = true;
        store_lv(V, s.pos, s.norm, s.emissive, mk3(0.f, 0.f, 0.f), s);
      }
    } else {
      TraceHit h;
      uint32_t nn = 0, nt = 0;
      if (traverse4<false>(sc, o, d, kInfF, (float)kBiasEps, false, lstk, h, nn, nt)) {   // :159-299
        const float w = (1.f - h.u) - h.v;
        const Surf s = fetch_surf(sc, h.leaf, w, h.u, h.v, true);
        const f3 nl = dot(s.norm, d) < 0.f ? s.norm : s.norm * -1.f;
        const Bounce b = sample_bounce(st, inc, o, d, h.t, s, nl, RT_LIGHT, false, true);
        no = b.orig;
        nd = b.dir;
        nvalid = true;
        mulT = vmul(b.mul, irr);
        store_lv(V, s.pos, s.norm, irr, d * -1.f, s);
      }
    }
    valid = nvalid;
    o = no;
    d = nd;
    irr = mulT;
  }
}

template <int STK, bool Q>
__global__ __launch_bounds__(kBlock) void k_regen_eye(DevScene sc, FrameArgs fa, const LightVert* __restrict__ lv,
                                                      float* __restrict__ result, uint32_t* __restrict__ sample_n,
                                                      DevStats* __restrict__ stats) {
  __shared__ uint64_t lstk_all[STK * kBlock];
  LdsU64* lstk = lds_column(lstk_all);
  const uint32_t W = fa.width, H = fa.height, pix = blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t ext = 0, shr = 0, paths = 0, en = 0, et = 0, sn = 0, stc = 0;
  if (pix < W * H) {
    const uint32_t levels = fa.eye_depth < kRegenMaxLevels ? fa.eye_depth : kRegenMaxLevels;
    const uint32_t y = pix / W, x = pix - y * W;
    const uint64_t inc = ((uint64_t)pix << 1) | 1u;
    uint64_t st;
    pcg32_seed(st, inc, stream_key(fa.seed, fa.frame0, kDomainRegenEye, 0u));
    const f3 cpos = mk3(fa.cam_pos[0], fa.cam_pos[1], fa.cam_pos[2]);
    const f3 cright = mk3(fa.cam_right[0], fa.cam_right[1], fa.cam_right[2]);
    const f3 cup = mk3(fa.cam_up[0], fa.cam_up[1], fa.cam_up[2]);
    const f3 cdir = mk3(fa.cam_dir[0], fa.cam_dir[1], fa.cam_dir[2]);
    const float Wf = (float)W, Hf = (float)H, tf = fa.tan_half;
    const float pu = ((((2.f * ((float)x + 0.5f)) / Wf - 1.f) * tf) * Wf) / Hf;   // pt0_kernel :646-647
    const float pv = ((2.f * ((float)y + 0.5f)) / Hf - 1.f) * tf;
    auto camera_ray = [&]() {
      const float au = pu + ((uniform01(st, inc) - 0.5f) / Hf) * tf;
      const float av = pv + ((uniform01(st, inc) - 0.5f) / Hf) * tf;
      return normalize((cright * au + cup * av) + cdir);
    };
    f3 o = cpos, d = camera_ray();
    float weight = 1.f;
    uint32_t depth = 0, count = 0, D = levels;
    f3 mul[kRegenMaxLevels], em[kRegenMaxLevels], con[kRegenMaxLevels];   // per level: f, emission, connection
    uint32_t bank = 0;                                     // levels where a camera path starts
    paths = 1;
    for (uint32_t L = 0; L < levels; L++) {
      TraceHit h;
      ext++;
      if (!traverse4<false, STK, Q>(sc, o, d, kInfF, (float)kBiasEps, false, lstk, h, en, et)) {   // :620-623
        D = L;
        break;
      }
      const float w = (1.f - h.u) - h.v;
      const Surf s = fetch_surf(sc, h.leaf, w, h.u, h.v, true);
      const f3 nl = dot(s.norm, d) < 0.f ? s.norm : s.norm * -1.f;
      const Bounce b = sample_bounce(st, inc, o, d, h.t, s, nl, RT_EYE, false, true);
      weight = weight * length(b.mul);
      if (length(s.emissive) > 0.f) weight = 0.f;
      if (depth == 0) bank |= 1u << L;
      f3 c = mk3(0.f, 0.f, 0.f), no = b.orig, nd = b.dir;
      if (!prob_rand(st, inc, weight)) {                 // a new camera path for the pixel (:562-571)
        weight = 1.f;
        depth = 0;
        count++;
        no = cpos;
        nd = camera_ray();
        paths++;
      } else {
        depth++;
        if (b.type == RT_DIFF) {                          // GetLightFromRandLightVertices (:400-431)
          const f3 P = s.pos + nl * (float)kBiasEps;
          uint32_t li = (uint32_t)(uniform01(st, inc) * (float)kRegenLightN);
          if (li >= kRegenLightN) li = kRegenLightN - 1;
          const LightVert& LV = lv[li];
          const f3 lpos = mk3(LV.pos[0], LV.pos[1], LV.pos[2]), lirr = mk3(LV.irrad[0], LV.irrad[1], LV.irrad[2]);
          const f3 ldir = mk3(LV.irradDir[0], LV.irradDir[1], LV.irradDir[2]);
          f3 irrad = mk3(0.f, 0.f, 0.f), irradDir = irrad;
          const f3 toL = normalize(lpos - P);
          const float dist = length(lpos - P);
          if (length(lirr) > 0.f && dot(nl, toL) > 0.f) {
            TraceHit hs;
            shr++;
            if (!traverse4<true, STK, Q>(sc, P, toL, (float)shadow_max((double)dist, fa.flags), (float)kBiasEps,
                                         false, lstk, hs, sn, stc)) {
              if (length(ldir) > (float)kFltEps)
                irrad = shading(toL * -1.f, ldir, lirr, mk3(LV.norm[0], LV.norm[1], LV.norm[2]),
                                mk3(LV.diff[0], LV.diff[1], LV.diff[2]), LV.metallic, LV.roughness, LV.specular,
                                1.f, 1.f) +
                        mk3(LV.emissive[0], LV.emissive[1], LV.emissive[2]);
              else
                irrad = lirr;
              irradDir = toL;
            }
          }
          c = vmax0(shading(d * -1.f, irradDir, irrad, nl, s.diff, s.metal, s.rough, s.spec, 1.f - s.trans, 1.f));
          if (length(c) > 0.f) count += 4;
        }
      }
      mul[L] = b.mul;
      em[L] = s.emissive;
      con[L] = c;
      o = no;
      d = nd;
    }
    count++;                                               // the miss or the level bound (:620-630)
    f3 r = mk3(0.f, 0.f, 0.f), accum = r;
    for (uint32_t k = D; k-- > 0;) {
      r = (vmul(mul[k], r) + em[k]) + con[k];
      if ((bank >> k) & 1u) {
        accum = accum + r;
        r = mk3(0.f, 0.f, 0.f);
      }
    }
    const uint32_t nold = fa.frame0 == 0 ? 0u : sample_n[pix], nn = nold + count;
    const float ri = 1.f / (float)nn, oi = (float)nold * ri;
    float* res = result + (size_t)pix * 3;
    res[0] = fmaxf(ri * accum.x + oi * res[0], 0.f);
    res[1] = fmaxf(ri * accum.y + oi * res[1], 0.f);
    res[2] = fmaxf(ri * accum.z + oi * res[2], 0.f);
    sample_n[pix] = nn;
  }
  if (stats) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats->frames, 1ull);
    const uint32_t a0 = wave_sum(ext), a1 = wave_sum(shr), a2 = wave_sum(paths);
    const uint32_t b0 = wave_sum(en), b1 = wave_sum(et), b2 = wave_sum(sn), b3 = wave_sum(stc);
    if (lane == 0 && (a0 | a2)) {
      atomicAdd(&stats->ext_rays, (unsigned long long)a0);
      atomicAdd(&stats->shadow_rays, (unsigned long long)a1);
      atomicAdd(&stats->eye_paths, (unsigned long long)a2);
      atomicAdd(&stats->ext_nodes, (unsigned long long)b0);
      atomicAdd(&stats->ext_tris, (unsigned long long)b1);
      atomicAdd(&stats->sh_nodes, (unsigned long long)b2);
      atomicAdd(&stats->sh_tris, (unsigned long long)b3);
    }
  }
}

hipError_t launch_regen_light(const DevScene& sc, const FrameArgs& fa, LightVert* lv, hipStream_t st) {
  if (!lv || !sc.tris || (sc.n_light_tris && !sc.light_tris)) return hipErrorInvalidValue;
  k_regen_light<<<1, kRegenLightN / kRegenLightDepth, 0, st>>>(sc, fa, lv);
  return hipGetLastError();
}
hipError_t launch_regen_eye(const DevScene& sc, const FrameArgs& fa, const LightVert* lv, float* result,
                            uint32_t* sample_n, int lds_stack, int quant, DevStats* stats, hipStream_t st) {
  const uint64_t n = (uint64_t)fa.width * fa.height;
  if (n == 0) return hipSuccess;
  if (!lv || !result || !sample_n || n > 0xFFFFFFFFull - kBlock || fa.eye_depth < 1 ||
      fa.eye_depth > kRegenMaxLevels || !sc.tris || !(quant ? (const void*)sc.nodes4q : (const void*)sc.nodes4))
    return hipErrorInvalidValue;
  const uint32_t grid = (uint32_t)((n + kBlock - 1) / kBlock);
#define CPT_RG(S, W, Q) k_regen_eye<S, Q><<<grid, kBlock, 0, st>>>(sc, fa, lv, result, sample_n, stats)
#define CPT_RG_Q(S, W) CPT_RG(S, W, true)
#define CPT_RG_F(S, W) CPT_RG(S, W, false)
  if (quant) {
    CPT_VARIANT_SWITCH(lds_stack, CPT_RG_Q)
  } else {
    CPT_VARIANT_SWITCH(lds_stack, CPT_RG_F)
  }
#undef CPT_RG_Q
#undef CPT_RG_F
#undef CPT_RG
  return hipGetLastError();
}

// ---------------------------------------------------------------- megakernel
// The reference's regeneration baseline (cudaRTBDPTRegen: one thread carries
// its sample through every bounce, src/cudaRTBDPTRegen.cu) over this library's
// functions: one thread per sample slot of the batch generates the eye path
// (eye_path), and per bounce traces the closest hit (traverse4), shades it
// (shade_path) and traces the connection's occlusion ray, adding the
// contribution when nothing occludes it.  The per-path arithmetic and its order
// are those of the wavefront passes (k_gen_eye, k_trace_dyn, k_shade,
// k_shadow), so the image is the same bit for bit; only the schedule differs:
// no path buffers, no compaction, and lanes of a wave idle once their paths
// end.  The frame counters receive the same per-bounce counts.
// One bounce of a path that one thread carries (k_mega, k_tail): the closest
// hit, shade_path, and the occlusion ray of the connection it sets up, whose
// contribution is added when nothing occludes it.
template <int STK, bool Q>
__device__ __forceinline__ void carry_bounce(const DevScene& sc, const FrameArgs& fa, FrameCounters* fc,
                                             const LightVert* lv, const Accum& ac, PathState& p, LdsU64* lstk,
                                             bool& survive, bool& emit, uint32_t& en, uint32_t& et, uint32_t& sn,
                                             uint32_t& st) {
  TraceHit h;
  const bool hit = traverse4<false, STK, Q>(sc, ld3(p.ro), ld3(p.rd), p.ro.w, p.rd.w, false, lstk, h, en, et);
  const float4 hv = make_float4(hit ? h.t : kInfF, __uint_as_float(hit ? h.leaf : kInvalid), h.u, h.v);
  ShadowRay sr;
  sr.o = sr.d = sr.c = make_float4(0.f, 0.f, 0.f, 0.f);
  shade_path(sc, fa, fc, lv, ac, p, hv, survive, emit, sr);
  if (emit) {
    TraceHit hs;
    if (!traverse4<true, STK, Q>(sc, ld3(sr.o), ld3(sr.d), sr.o.w, sr.d.w, false, lstk, hs, sn, st)) {
      p.acc.x = p.acc.x + sr.c.x;
      p.acc.y = p.acc.y + sr.c.y;
      p.acc.z = p.acc.z + sr.c.z;
      p.acc.w = __uint_as_float(__float_as_uint(p.acc.w) + 4u);
    }
  }
}

// Per-bounce live and shadow counts of a workgroup into the frame counters,
// and its traversal counters into the statistics (TAIL: the tail's own
// counters, with nr extension rays).
template <bool TAIL>
__device__ __forceinline__ void carry_flush(FrameCounters* fc, DevStats* stats, const uint32_t* n_live,
                                            const uint32_t* n_shadow, uint32_t depth, uint32_t en, uint32_t et,
                                            uint32_t sn, uint32_t st, uint32_t nr = 0) {
  if (threadIdx.x < depth) {
    const uint32_t b = threadIdx.x;
    if (b + 1 < 16 && n_live[b]) atomicAdd(&fc->active[b + 1], n_live[b]);
    if (n_shadow[b]) atomicAdd(&fc->shadow[b], n_shadow[b]);
  }
  if (stats && TAIL) {
    const uint32_t a0 = wave_sum(en + sn), a1 = wave_sum(et + st), a2 = wave_sum(nr);
    if ((threadIdx.x & 63) == 0 && (a0 | a1 | a2)) {
      atomicAdd(&stats->tail_nodes, (unsigned long long)a0);
      atomicAdd(&stats->tail_tris, (unsigned long long)a1);
      atomicAdd(&stats->tail_rays, (unsigned long long)a2);
    }
  } else if (stats) {
    const uint32_t a0 = wave_sum(en), a1 = wave_sum(et), a2 = wave_sum(sn), a3 = wave_sum(st);
    if ((threadIdx.x & 63) == 0) {
      if (a0 | a1) {
        atomicAdd(&stats->ext_nodes, (unsigned long long)a0);
        atomicAdd(&stats->ext_tris, (unsigned long long)a1);
      }
      if (a2 | a3) {
        atomicAdd(&stats->sh_nodes, (unsigned long long)a2);
        atomicAdd(&stats->sh_tris, (unsigned long long)a3);
      }
    }
  }
}

template <int STK, bool Q>
__global__ __launch_bounds__(kBlock) void k_mega(DevScene sc, FrameArgs fa, FrameCounters* __restrict__ fc,
                                                 const LightVert* __restrict__ lv, Accum ac,
                                                 DevStats* __restrict__ stats) {
  __shared__ uint64_t lstk_all[STK * kBlock];
  __shared__ uint32_t n_live[16], n_shadow[16];
  LdsU64* lstk = lds_column(lstk_all);
  if (threadIdx.x < 16) { n_live[threadIdx.x] = 0u; n_shadow[threadIdx.x] = 0u; }
  __syncthreads();
  const uint32_t wh = fa.width * fa.height;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t en = 0, et = 0, sn = 0, st = 0;
  bool live = i < wh * fa.nframes;
  PathState p;
  if (live) {                                               // pixel-major, as k_gen_eye
    const uint32_t K = fa.nframes, pix = i / K, k = i - pix * K;
    p = eye_path(fa, pix, k * wh + pix, stream_key(fa.seed, frame_base(fa) + k, 1u, 0u));
    p.fk = k;
  }
  const uint32_t depth = fa.eye_depth < 16u ? fa.eye_depth : 16u;
  for (uint32_t b = 0; b < depth; b++) {
    if (__ballot(live) == 0ull) break;                    // wave-uniform
    bool survive = false, emit = false;
    if (live) {
      carry_bounce<STK, Q>(sc, fa, fc, lv, ac, p, lstk, survive, emit, en, et, sn, st);
      live = survive;
    }
    const uint32_t cs = (uint32_t)__popcll(__ballot(survive)), ce = (uint32_t)__popcll(__ballot(emit));
    if (lane == 0) {
      if (cs) atomicAdd(&n_live[b], cs);
      if (ce) atomicAdd(&n_shadow[b], ce);
    }
  }
  __syncthreads();
  carry_flush<false>(fc, stats, n_live, n_shadow, depth, en, et, sn, st);
}

// The last bounces of the wavefront in one launch.  Once few paths are left,
// a bounce's refilling grid runs partly empty and each of its three launches
// (extension, shading, occlusion) waits for its slowest ray; from bounce
// `bounce0` on, each thread instead takes live slots of `cur` (strided over a
// persistent grid) and carries each path through the remaining bounces as
// k_mega does, with the same per-path arithmetic: the image does not change.
template <int STK, bool Q>
__global__ __launch_bounds__(kBlock) void k_tail(DevScene sc, FrameArgs fa, PathBuf cur,
                                                 FrameCounters* __restrict__ fc, uint32_t bounce0,
                                                 const LightVert* __restrict__ lv, Accum ac,
                                                 DevStats* __restrict__ stats) {
  __shared__ uint64_t lstk_all[STK * kBlock];
  __shared__ uint32_t n_live[16], n_shadow[16];
  LdsU64* lstk = lds_column(lstk_all);
  if (threadIdx.x < 16) { n_live[threadIdx.x] = 0u; n_shadow[threadIdx.x] = 0u; }
  __syncthreads();
  const uint32_t count = fc->active[bounce0];
  const int lane = threadIdx.x & 63;
  const uint32_t depth = fa.eye_depth < 16u ? fa.eye_depth : 16u;
  uint32_t en = 0, et = 0, sn = 0, st = 0, nr = 0;
  for (uint32_t base = blockIdx.x * kBlock; base < count; base += gridDim.x * kBlock) {   // workgroup-uniform
    const uint32_t i = base + threadIdx.x;
    bool live = i < count;
    PathState p;
    if (live) {
      p.ro = cur.ray_o[i];
      p.rd = cur.ray_d[i];
      p.pix = cur.pix[i];
      p.sid = cur.sid[i];
      p.rng = cur.rng[i];
      const uint32_t dk = cur.depth[i];
      unpack_dk(p, dk);
      if (dk & kFreshPath) {
        p.mul = make_float4(1.f, 1.f, 1.f, 0.f);
        p.smp = make_float4(0.f, 0.f, 0.f, 0.f);
        p.acc = p.smp;
      } else {
        p.mul = cur.mul[i];
        p.smp = cur.smp[i];
        p.acc = cur.acc[i];
      }
    }
    for (uint32_t b = bounce0; b < depth; b++) {
      if (__ballot(live) == 0ull) break;                  // wave-uniform
      bool survive = false, emit = false;
      if (live) {
        carry_bounce<STK, Q>(sc, fa, fc, lv, ac, p, lstk, survive, emit, en, et, sn, st);
        live = survive;
        nr++;
      }
      const uint32_t cs = (uint32_t)__popcll(__ballot(survive)), ce = (uint32_t)__popcll(__ballot(emit));
      if (lane == 0) {
        if (cs) atomicAdd(&n_live[b], cs);
        if (ce) atomicAdd(&n_shadow[b], ce);
      }
    }
  }
  __syncthreads();
  carry_flush<true>(fc, stats, n_live, n_shadow, depth, en, et, sn, st, nr);
}
hipError_t launch_tail(const DevScene& sc, const FrameArgs& fa, PathBuf cur, FrameCounters* fc, uint32_t bounce0,
                       uint32_t grid, const LightVert* lv, const Accum& ac, int lds_stack, int quant,
                       DevStats* stats, hipStream_t st) {
  // bounce0 >= 1: no path of the tail is at its first hit, so a G-buffer
  // variant (written at depth 0 only) is allowed
  if (!fc || !lv || !ac.rec.acc || ac.apply != 2 || fa.eye_depth > 16 ||
      bounce0 == 0 || bounce0 >= fa.eye_depth || grid == 0 || !pathbuf_ok(cur) || !sc.tris ||
      !(quant ? (const void*)sc.nodes4q : (const void*)sc.nodes4))
    return hipErrorInvalidValue;
#define CPT_TL(S, W, Q) k_tail<S, Q><<<grid, kBlock, 0, st>>>(sc, fa, cur, fc, bounce0, lv, ac, stats)
#define CPT_TL_Q(S, W) CPT_TL(S, W, true)
#define CPT_TL_F(S, W) CPT_TL(S, W, false)
  if (quant) {
    CPT_VARIANT_SWITCH(lds_stack, CPT_TL_Q)
  } else {
    CPT_VARIANT_SWITCH(lds_stack, CPT_TL_F)
  }
#undef CPT_TL_Q
#undef CPT_TL_F
#undef CPT_TL
  return hipGetLastError();
}
hipError_t launch_mega(const DevScene& sc, const FrameArgs& fa, FrameCounters* fc, const LightVert* lv,
                       const Accum& ac, int lds_stack, int quant, DevStats* stats, hipStream_t st) {
  const uint64_t n = (uint64_t)fa.width * fa.height * fa.nframes;
  if (n == 0) return hipSuccess;
  if (n > 0xFFFFFFFFull - kBlock || !fc || !lv || !ac.rec.acc || ac.apply != 2 || (ac.variant & kVariantGbuf) ||
      fa.eye_depth > 16 || !sc.tris ||
      !(quant ? (const void*)sc.nodes4q : (const void*)sc.nodes4))
    return hipErrorInvalidValue;
  dim3 grid((uint32_t)((n + kBlock - 1) / kBlock)), block(kBlock);
#define CPT_MG(S, W, Q) k_mega<S, Q><<<grid, block, 0, st>>>(sc, fa, fc, lv, ac, stats)
  if (quant) {
#define CPT_MG_Q(S, W) CPT_MG(S, W, true)
#define CPT_MG_F(S, W) CPT_MG(S, W, false)
    CPT_VARIANT_SWITCH(lds_stack, CPT_MG_Q)
  } else {
    CPT_VARIANT_SWITCH(lds_stack, CPT_MG_F)
  }
#undef CPT_MG_Q
#undef CPT_MG_F
#undef CPT_MG
  return hipGetLastError();
}

// ---------------------------------------------------------------- adaptive sampling
// e^x: Cody-Waite reduction + degree-7 Taylor polynomial, 2^k by ldexp
// (Cody & Waite 1980).  The host oracle evaluates the identical fma sequence.
__device__ __forceinline__ float cpt_expf(float x) {
  if (x < -103.0f) return 0.f;
  if (x > 88.0f) return __builtin_huge_valf();
  float k = rintf(x * 1.44269504f);
  float r = fmaf(k, -0.693145751953125f, x);
  r = fmaf(k, -1.42860676533018e-06f, r);
  float p = 1.f / 5040.f;
  p = fmaf(p, r, 1.f / 720.f);
  p = fmaf(p, r, 1.f / 120.f);
  p = fmaf(p, r, 1.f / 24.f);
  p = fmaf(p, r, 1.f / 6.f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.f);
  p = fmaf(p, r, 1.f);
  return ldexpf(p, (int)k);
}
__device__ __forceinline__ float sqdist3(const float* a, const float* b) {
  float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return (dx * dx + dy * dy) + dz * dz;
}
// |a - b|^2 of a filter range term: fma(dz, dz, fma(dy, dy, dx * dx)).
__device__ __forceinline__ float sqd_fma(const float* a, const float* b) {
  const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}
// Coordinate as the reference clamps it: (x + j) is unsigned, so a negative
// coordinate wraps and lands on the high edge (cudaBilateralFilter.cu:85).
__device__ __forceinline__ uint32_t ref_clamp_i(int v, uint32_t n) {
  return (v < 0 || (uint32_t)v >= n) ? n - 1 : (uint32_t)v;
}

constexpr int kTile = 16;
constexpr int kLdsRadius = 9;     // LDS tile path up to this radius, global reads beyond

// Cross-bilateral over colour/position/normal(/albedo) + the weighted squared
// residual of pt_calculateSquareError_kernel, one 16x16 tile per workgroup.
// The (16+2r)^2 input halo is staged in LDS once; every tap then reads LDS.
// One exponential per tap: the range terms' exponents are summed, each scaled
// by its 1 / 2 sigma^2 (correctly rounded once per thread) and accumulated
// with fmas, so a tap takes no division.
template <bool DIFF, bool LDS>
__global__ __launch_bounds__(kTile * kTile) void k_filter_error(FilterArgs fa, const float* __restrict__ color,
                                                                const float* __restrict__ gpos,
                                                                const float* __restrict__ gnrm,
                                                                const float* __restrict__ gdif,
                                                                const float* __restrict__ gauss,
                                                                const uint32_t* __restrict__ sample_n,
                                                                float* __restrict__ filtered, float* __restrict__ err) {
  extern __shared__ float sm[];
  const int R = (int)fa.radius, TW = kTile + 2 * R, TA = TW * TW;
  const uint32_t W = fa.width, H = fa.height;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * kTile + tx;
  const int bx = blockIdx.x * kTile, by = blockIdx.y * kTile;
  float* sg = sm;                                  // 2R+1 spatial weights
  float* sc = sg + 2 * kMaxFilterRadius + 2;       // 3 floats per halo pixel, per plane
  float* sp = sc + 3 * (LDS ? TA : 0);
  float* sn = sp + 3 * (LDS ? TA : 0);
  float* sd = sn + 3 * (LDS ? TA : 0);
  float* sz = sd + 3 * (LDS && DIFF ? TA : 0);     // depth per halo pixel (depth term on)
  const bool depth = fa.kz > 0.f;
  const float cam[3] = {fa.cam[0], fa.cam[1], fa.cam[2]};
  for (int k = tid; k < 2 * R + 1; k += kTile * kTile) sg[k] = gauss[k];
  if (LDS) {
    for (int k = tid; k < TA; k += kTile * kTile) {
      int hx = k % TW, hy = k / TW;
      size_t g = ((size_t)ref_clamp_i(by + hy - R, H) * W + ref_clamp_i(bx + hx - R, W)) * 3;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        sc[k * 3 + c] = color[g + c];
        sp[k * 3 + c] = gpos[g + c];
        sn[k * 3 + c] = gnrm[g + c];
        if (DIFF) sd[k * 3 + c] = gdif[g + c];
      }
      if (depth) sz[k] = sqrtf(sqdist3(gpos + g, cam));
    }
  }
  __syncthreads();
  const uint32_t x = bx + tx, y = by + ty;
  if (x >= W || y >= H) return;
  const size_t ind = (size_t)y * W + x;
  float cc[3], cp[3], cn[3], cd[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    cc[c] = color[ind * 3 + c];
    cp[c] = gpos[ind * 3 + c];
    cn[c] = gnrm[ind * 3 + c];
    cd[c] = DIFF ? gdif[ind * 3 + c] : 0.f;
  }
  const float zc = depth ? sqrtf(sqdist3(cp, cam)) : 0.f;
  const float izc = zc > 0.f ? 1.f / zc : 0.f;
  const float ikc = 1.f / fa.kc, ikp = 1.f / fa.kp, ikn = 1.f / fa.kn, ikd = 1.f / fa.kd;
  const float ikz = depth ? 1.f / fa.kz : 0.f;
  float sum = 0.f, t0 = 0.f, t1 = 0.f, t2 = 0.f;
  for (int i = -R; i <= R; i++) {
    for (int j = -R; j <= R; j++) {
      const float *qc, *qp, *qn, *qd = nullptr;
      float zq = 0.f;
      if (LDS) {
        int k = (ty + R + i) * TW + (tx + R + j);
        qc = sc + k * 3; qp = sp + k * 3; qn = sn + k * 3;
        if (DIFF) qd = sd + k * 3;
        if (depth) zq = sz[k];
      } else {
        size_t g = ((size_t)ref_clamp_i((int)y + i, H) * W + ref_clamp_i((int)x + j, W)) * 3;
        qc = color + g; qp = gpos + g; qn = gnrm + g;
        if (DIFF) qd = gdif + g;
        if (depth) zq = sqrtf(sqdist3(qp, cam));
      }
      float a = sqd_fma(qc, cc) * ikc;
      a = __builtin_fmaf(sqd_fma(qp, cp), ikp, a);
      a = __builtin_fmaf(sqd_fma(qn, cn), ikn, a);
      if (DIFF) a = __builtin_fmaf(sqd_fma(qd, cd), ikd, a);
      if (depth) {
        const float dz = (zq - zc) * izc;
        a = __builtin_fmaf(dz * dz, ikz, a);
      }
      float f = (sg[i + R] * sg[j + R]) * cpt_expf(-a);
      t0 = __builtin_fmaf(f, qc[0], t0);
      t1 = __builtin_fmaf(f, qc[1], t1);
      t2 = __builtin_fmaf(f, qc[2], t2);
      sum = sum + f;
    }
  }
  float o0 = t0 / sum, o1 = t1 / sum, o2 = t2 / sum;
  filtered[ind * 3] = o0;
  filtered[ind * 3 + 1] = o1;
  filtered[ind * 3 + 2] = o2;
  // pt_calculateSquareError_kernel (AdapBilatAtomic.cu:1143-1156)
  uint32_t sn_ = sample_n[ind];
  float red = (float)fa.nd / (float)(fa.nd + sn_);
  if (fa.nd + sn_ < sn_) red = 0.f;
  float d0 = o0 - cc[0], d1 = o1 - cc[1], d2 = o2 - cc[2];
  err[ind] = red * ((d0 * d0 + d1 * d1) + d2 * d2) / 3.f;
}

hipError_t launch_filter_error(const FilterArgs& fa, const float* color, const float* gpos, const float* gnrm,
                               const float* gdif, const float* gauss, const uint32_t* sample_n, float* filtered,
                               float* err, hipStream_t st) {
  if (!color || !gpos || !gnrm || !gauss || !sample_n || !filtered || !err || (fa.use_diff && !gdif) ||
      (int)fa.radius > kMaxFilterRadius || fa.width == 0 || fa.height == 0)
    return hipErrorInvalidValue;
  dim3 block(kTile, kTile), grid((fa.width + kTile - 1) / kTile, (fa.height + kTile - 1) / kTile);
  const bool lds = (int)fa.radius <= kLdsRadius;
  const int TW = kTile + 2 * (int)fa.radius;
  size_t bytes = sizeof(float) * (2 * kMaxFilterRadius + 2);
  if (lds) bytes += sizeof(float) * TW * TW * (3 * (fa.use_diff ? 4 : 3) + (fa.kz > 0.f ? 1 : 0));
  if (fa.use_diff) {
    if (lds) k_filter_error<true, true><<<grid, block, bytes, st>>>(fa, color, gpos, gnrm, gdif, gauss, sample_n, filtered, err);
    else k_filter_error<true, false><<<grid, block, bytes, st>>>(fa, color, gpos, gnrm, gdif, gauss, sample_n, filtered, err);
  } else {
    if (lds) k_filter_error<false, true><<<grid, block, bytes, st>>>(fa, color, gpos, gnrm, gdif, gauss, sample_n, filtered, err);
    else k_filter_error<false, false><<<grid, block, bytes, st>>>(fa, color, gpos, gnrm, gdif, gauss, sample_n, filtered, err);
  }
  return hipGetLastError();
}

// Fixed-order double sum (stands in for thrust::reduce, AdapBilatAtomic.cu:1444):
// a pairwise tree inside each 256-element block; the same tree over each chunk
// of 256 block sums; the chunk sums in order.
__global__ __launch_bounds__(kBlock) void k_err_block_sums(const float* __restrict__ e, uint32_t n,
                                                           double* __restrict__ bs) {
  __shared__ double s[kBlock];
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  s[threadIdx.x] = i < n ? (double)e[i] : 0.0;
  __syncthreads();
  for (int stp = kBlock / 2; stp >= 1; stp >>= 1) {
    if ((int)threadIdx.x < stp) s[threadIdx.x] += s[threadIdx.x + stp];
    __syncthreads();
  }
  if (threadIdx.x == 0) bs[blockIdx.x] = s[0];
}
// One workgroup of four 256-thread groups, each reducing one chunk of block
// sums at a time (their loads in parallel, where one thread walking the block
// sums waits on every load in turn); the chunk sums wait in LDS for thread 0.
constexpr uint32_t kErrChunks = 1024;   // 2^26 pixels: 262,144 block sums
__global__ __launch_bounds__(1024) void k_err_final(const double* __restrict__ bs, uint32_t nb,
                                                    float* __restrict__ sum) {
  __shared__ double s[1024];
  __shared__ double cs[kErrChunks];
  const uint32_t g = threadIdx.x >> 8, t = threadIdx.x & 255u;
  const uint32_t nc = (nb + 255u) / 256u;
  for (uint32_t c0 = 0; c0 < nc; c0 += 4) {
    const uint32_t c = c0 + g, i = c * 256u + t;
    s[threadIdx.x] = (c < nc && i < nb) ? bs[i] : 0.0;
    __syncthreads();
    for (uint32_t stp = 128; stp >= 1; stp >>= 1) {
      if (t < stp) s[threadIdx.x] += s[threadIdx.x + stp];
      __syncthreads();
    }
    if (t == 0 && c < nc) cs[c] = s[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (uint32_t c = 0; c < nc; c++) tot += cs[c];
    *sum = (float)tot;
  }
}
hipError_t launch_error_sum(const float* err, uint32_t n, double* block_sums, float* sum_out, hipStream_t st) {
  uint32_t nb = (n + kBlock - 1) / kBlock;
  if (n == 0 || (nb + 255u) / 256u > kErrChunks) return hipErrorInvalidValue;
  k_err_block_sums<<<nb, kBlock, 0, st>>>(err, n, block_sums);
  k_err_final<<<1, 1024, 0, st>>>(block_sums, nb, sum_out);
  return hipGetLastError();
}

// pt_genAdapPathQueue_kernel's per-pixel count (AdapBilatAtomic.cu:830-831).
__device__ __forceinline__ uint32_t prop_count(float e, float sum, uint32_t gen) {
  float r = (float)gen * (e / sum);
  if (!(r >= 0.f) || !(r < 4294967040.f)) return 0u;
  return (uint32_t)r;
}

constexpr int kScanPer = 4;                       // elements per thread in the scan
constexpr int kScanBlock = kBlock * kScanPer;     // 1024 elements per workgroup

// Counts and workgroup-local exclusive offsets.
// With a per-pixel cap (maxp > 0) a count stops at maxp - paths.
__device__ __forceinline__ uint32_t cap_count(uint32_t c, const uint32_t* __restrict__ paths, uint32_t maxp,
                                              uint32_t i) {
  if (maxp == 0u) return c;
  const uint32_t have = paths[i];
  const uint32_t room = have < maxp ? maxp - have : 0u;
  return c < room ? c : room;
}

__global__ __launch_bounds__(kBlock) void k_scan_local(const float* __restrict__ err, const float* __restrict__ sum,
                                                       uint32_t n, uint32_t gen, uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ offsets, uint32_t* __restrict__ btot,
                                                       const uint32_t* __restrict__ paths, uint32_t maxp) {
  __shared__ uint32_t s[kBlock];
  const float sm = *sum;
  const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
  uint32_t c[kScanPer], tsum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) {
    uint32_t i = base + k;
    c[k] = i < n ? cap_count(prop_count(err[i], sm, gen), paths, maxp, i) : 0u;
    tsum += c[k];
  }
  s[threadIdx.x] = tsum;
  __syncthreads();
  for (int o = 1; o < kBlock; o <<= 1) {          // Hillis-Steele inclusive scan of thread sums
    uint32_t v = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tsum;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) {
    uint32_t i = base + k;
    if (i < n) { counts[i] = c[k]; offsets[i] = run; }
    run += c[k];
  }
  if (threadIdx.x == kBlock - 1) btot[blockIdx.x] = s[kBlock - 1];
}
// Exclusive scan of the workgroup totals (one workgroup, chunks of 256).
__global__ __launch_bounds__(kBlock) void k_scan_totals(uint32_t* __restrict__ btot, uint32_t nb,
                                                        uint32_t* __restrict__ total) {
  __shared__ uint32_t s[kBlock];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < nb; c0 += kBlock) {
    uint32_t i = c0 + threadIdx.x;
    uint32_t v = i < nb ? btot[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
      uint32_t w = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
      __syncthreads();
      s[threadIdx.x] += w;
      __syncthreads();
    }
    if (i < nb) btot[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += s[kBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}
__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* __restrict__ offsets, const uint32_t* __restrict__ btot,
                                                     uint32_t n) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) offsets[i] += btot[i / kScanBlock];
}
hipError_t launch_alloc_scan(const float* err, const float* sum, uint32_t n, uint32_t gen, uint32_t* counts,
                             uint32_t* offsets, uint32_t* block_tot, uint32_t* total, uint32_t cap,
                             const uint32_t* paths, uint32_t maxp, hipStream_t st) {
  (void)cap;
  if (maxp && !paths) return hipErrorInvalidValue;
  uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
  k_scan_local<<<nb, kBlock, 0, st>>>(err, sum, n, gen, counts, offsets, block_tot, paths, maxp);
  k_scan_totals<<<1, kBlock, 0, st>>>(block_tot, nb, total);
  k_scan_add<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(offsets, block_tot, n);
  return hipGetLastError();
}

// Multi-view adaptive batches scan each view's counts on their own: view v's
// offsets then move past the slots of views 0..v-1 (base: exclusive scan of
// the per-view totals, by k_scan_totals), so all views form one queue.
__global__ __launch_bounds__(kBlock) void k_offset_views(uint32_t* __restrict__ offsets, uint32_t n, FrameArgs fa,
                                                         const uint32_t* __restrict__ base) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) offsets[i] += base[view_of(fa, i)];
}
hipError_t launch_view_bases(uint32_t* view_totals, uint32_t n_views, uint32_t* total, hipStream_t st) {
  if (!view_totals || !total || n_views == 0) return hipErrorInvalidValue;
  k_scan_totals<<<1, kBlock, 0, st>>>(view_totals, n_views, total);
  return hipGetLastError();
}
hipError_t launch_offset_views(uint32_t* offsets, uint32_t n, const FrameArgs& fa, const uint32_t* base,
                               hipStream_t st) {
  if (!offsets || !base || !fa.views || fa.n_views == 0) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  k_offset_views<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(offsets, n, fa, base);
  return hipGetLastError();
}

// One eye path per queue slot: the pixel is the last one whose offset is
// <= slot (binary search), the sample index is slot - offset.
__global__ __launch_bounds__(kBlock) void k_gen_adaptive(FrameArgs fa, const uint32_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ total, uint32_t cap,
                                                         PathBuf pb, SampleRec rec) {
  const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
  uint32_t t = *total;
  if (t > cap) t = cap;
  if (q >= t) return;
  const uint32_t n = fa.width * fa.height;
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= q) lo = mid; else hi = mid;
  }
  const uint32_t k = q - offsets[lo];
  (void)counts;
  write_eye_path(fa, pb, q, lo, q, stream_key(fa.seed, fa.frame0, 1u, k));
  zero_rec(rec, q);
}
hipError_t launch_gen_adaptive(const FrameArgs& fa, const uint32_t* offsets, const uint32_t* counts,
                               const uint32_t* total, uint32_t cap, PathBuf pb, SampleRec rec, hipStream_t st) {
  if (!pathbuf_ok(pb) || !rec_ok(rec) || !offsets || !total) return hipErrorInvalidValue;
  k_gen_adaptive<<<(cap + kBlock - 1) / kBlock, kBlock, 0, st>>>(fa, offsets, counts, total, cap, pb, rec);
  return hipGetLastError();
}

// Running-mean update of one pixel from the sums of its new records.
__device__ __forceinline__ void accum_apply(uint32_t pix, uint32_t tN, f3 ts, f3 tp, f3 tn, f3 td, f3 tq,
                                            float* __restrict__ result, uint32_t* __restrict__ sample_n,
                                            float* __restrict__ gpos, float* __restrict__ gnrm,
                                            float* __restrict__ gdif, float* __restrict__ sq) {
  uint32_t nold = sample_n[pix], nn = tN + nold;
  if (nn > nold) {
    float ri = 1.f / (float)nn, oi = (float)nold * ri;
    float* r = result + (size_t)pix * 3;
    r[0] = fmaxf(ri * ts.x + oi * r[0], 0.f);
    r[1] = fmaxf(ri * ts.y + oi * r[1], 0.f);
    r[2] = fmaxf(ri * ts.z + oi * r[2], 0.f);
    sample_n[pix] = nold + tN;
    f3 vals[3] = {tp, tn, td};
    float* dst[3] = {gpos + (size_t)pix * 3, gnrm + (size_t)pix * 3, gdif + (size_t)pix * 3};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      dst[k][0] = ri * vals[k].x + oi * dst[k][0];
      dst[k][1] = ri * vals[k].y + oi * dst[k][1];
      dst[k][2] = ri * vals[k].z + oi * dst[k][2];
    }
    float* q = sq + (size_t)pix * 3;                // squareResult (AdapEmp.cu:927-944)
    q[0] = fmaxf(ri * tq.x + oi * q[0], 0.f);
    q[1] = fmaxf(ri * tq.y + oi * q[1], 0.f);
    q[2] = fmaxf(ri * tq.z + oi * q[2], 0.f);
  }
}

// Adds record q to the running sums (records with N = 0 or an infinite
// component are skipped, AdapBilatAtomic.cu:1056-1060).
// One sample record: its four columns are loaded together (one round trip),
// then added in record order, so that loads of several records can be in
// flight while the sums keep their order.
struct RecVals {
  float4 a, p, n, d;
};
__device__ __forceinline__ RecVals load_record(const SampleRec& rec, uint32_t q) {
  RecVals r;
  r.a = rec.acc[q];
  r.p = rec.gpos[q];
  r.n = rec.gnrm[q];
  r.d = rec.gdif[q];
  return r;
}
__device__ __forceinline__ void add_record(const RecVals& r, f3& ts, f3& tp, f3& tn, f3& td, f3& tq, uint32_t& tN) {
  uint32_t N = __float_as_uint(r.a.w);
  f3 acc = ld3(r.a);
  if (N == 0 || isinf3(acc)) return;
  float nf = (float)N;
  ts = ts + acc;
  tq = tq + vmul(acc, acc);
  tN += N;
  tp = tp + ld3(r.p) * nf;
  tn = tn + ld3(r.n) * nf;
  td = td + ld3(r.d) * nf;
}
// Records q0, q0 + step, ... below q1 in order, four loads in flight.
__device__ __forceinline__ void accum_records(const SampleRec& rec, uint32_t q0, uint32_t q1, uint32_t step, f3& ts,
                                              f3& tp, f3& tn, f3& td, f3& tq, uint32_t& tN) {
  uint32_t q = q0;
  for (; q1 > 3 * step && q < q1 - 3 * step; q += 4 * step) {
    const RecVals r0 = load_record(rec, q), r1 = load_record(rec, q + step);
    const RecVals r2 = load_record(rec, q + 2 * step), r3 = load_record(rec, q + 3 * step);
    add_record(r0, ts, tp, tn, td, tq, tN);
    add_record(r1, ts, tp, tn, td, tq, tN);
    add_record(r2, ts, tp, tn, td, tq, tN);
    add_record(r3, ts, tp, tn, td, tq, tN);
  }
  for (; q < q1; q += step) add_record(load_record(rec, q), ts, tp, tn, td, tq, tN);
}

// Pixels with more records than this are summed by a whole workgroup.
constexpr uint32_t kAccumSerialMax = 32;


// pt_atomicApplyPathQueueResult_kernel + pt_accumTempResultToResult_kernel
// (AdapBilatAtomic.cu:1041-1141) with the per-pixel sums taken over the
// pixel's samples in queue order instead of float atomics.  Pixels holding
// more than kAccumSerialMax records are listed for k_accum_big.
__global__ __launch_bounds__(kBlock) void k_accum_adaptive(const uint32_t* __restrict__ offsets,
                                                           const uint32_t* __restrict__ counts, uint32_t n,
                                                           uint32_t cap, SampleRec rec, float* __restrict__ result,
                                                           uint32_t* __restrict__ sample_n, float* __restrict__ gpos,
                                                           float* __restrict__ gnrm, float* __restrict__ gdif,
                                                           float* __restrict__ sq, uint32_t* __restrict__ big_list,
                                                           uint32_t* __restrict__ big_n) {
  const uint32_t pix = blockIdx.x * kBlock + threadIdx.x;
  if (pix >= n) return;
  uint32_t q0 = offsets[pix], q1 = q0 + counts[pix];
  if (q1 > cap) q1 = cap;
  if (q0 > q1) q0 = q1;
  if (q1 > q0 + kAccumSerialMax) {
    big_list[atomicAdd(big_n, 1u)] = pix;
    return;
  }
  f3 ts = mk3(0.f, 0.f, 0.f), tp = ts, tn = ts, td = ts, tq = ts;
  uint32_t tN = 0;
  accum_records(rec, q0, q1, 1u, ts, tp, tn, td, tq, tN);
  accum_apply(pix, tN, ts, tp, tn, td, tq, result, sample_n, gpos, gnrm, gdif, sq);
}

// Heavily sampled pixels: one workgroup per pixel, thread t sums records
// q0 + t, q0 + t + 256, ..., then a fixed butterfly within each wave and the
// four wave sums in order (deterministic for a given record count).
__global__ __launch_bounds__(kBlock) void k_accum_big(const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ counts, uint32_t cap, SampleRec rec,
                                                      float* __restrict__ result, uint32_t* __restrict__ sample_n,
                                                      float* __restrict__ gpos, float* __restrict__ gnrm,
                                                      float* __restrict__ gdif, float* __restrict__ sq,
                                                      const uint32_t* __restrict__ big_list,
                                                      const uint32_t* __restrict__ big_n) {
  constexpr int kV = 15;
  __shared__ float red[kBlock / 64][kV];
  __shared__ uint32_t redn[kBlock / 64];
  const uint32_t nb = *big_n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t i = blockIdx.x; i < nb; i += gridDim.x) {
    const uint32_t pix = big_list[i];
    uint32_t q0 = offsets[pix], q1 = q0 + counts[pix];
    if (q1 > cap) q1 = cap;
    f3 ts = mk3(0.f, 0.f, 0.f), tp = ts, tn = ts, td = ts, tq = ts;
    uint32_t tN = 0;
    accum_records(rec, q0 + threadIdx.x, q1, kBlock, ts, tp, tn, td, tq, tN);
    float v[kV] = {ts.x, ts.y, ts.z, tp.x, tp.y, tp.z, tn.x, tn.y, tn.z, td.x, td.y, td.z, tq.x, tq.y, tq.z};
#pragma unroll
    for (int k = 0; k < kV; k++)
      for (int o = 32; o >= 1; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    for (int o = 32; o >= 1; o >>= 1) tN += __shfl_xor(tN, o, 64);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < kV; k++) red[wave][k] = v[k];
      redn[wave] = tN;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float a[kV];
      uint32_t sN = 0;
#pragma unroll
      for (int k = 0; k < kV; k++) a[k] = 0.f;
      for (int w = 0; w < kBlock / 64; w++) {
#pragma unroll
        for (int k = 0; k < kV; k++) a[k] += red[w][k];
        sN += redn[w];
      }
      accum_apply(pix, sN, mk3(a[0], a[1], a[2]), mk3(a[3], a[4], a[5]), mk3(a[6], a[7], a[8]),
                  mk3(a[9], a[10], a[11]), mk3(a[12], a[13], a[14]), result, sample_n, gpos, gnrm, gdif, sq);
    }
    __syncthreads();
  }
}

hipError_t launch_accum_adaptive(const uint32_t* offsets, const uint32_t* counts, uint32_t n, uint32_t cap,
                                 SampleRec rec, float* result, uint32_t* sample_n, float* gpos, float* gnrm,
                                 float* gdif, float* sq, uint32_t* big_list, uint32_t* big_n, hipStream_t st) {
  if (!rec_ok(rec) || !offsets || !counts || !result || !sample_n || !gpos || !gnrm || !gdif || !sq || !big_list ||
      !big_n)
    return hipErrorInvalidValue;
  k_fill_u32<<<1, 1, 0, st>>>(big_n, 0u);
  k_accum_adaptive<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(offsets, counts, n, cap, rec, result, sample_n, gpos,
                                                                  gnrm, gdif, sq, big_list, big_n);
  k_accum_big<<<1024, kBlock, 0, st>>>(offsets, counts, cap, rec, result, sample_n, gpos, gnrm, gdif, sq, big_list,
                                       big_n);
  return hipGetLastError();
}

// ---- empirical error and round-based allocation -----------------------------

__global__ __launch_bounds__(kBlock) void k_emp_error(const float* __restrict__ r, const float* __restrict__ q,
                                                      const uint32_t* __restrict__ sample_n, uint32_t n,
                                                      float* __restrict__ err) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= n) return;
  const uint32_t N = sample_n[x];
  const float* a = r + (size_t)x * 3;
  const float* b = q + (size_t)x * 3;
  float e = ((b[0] - a[0] * a[0]) + (b[1] - a[1] * a[1]) + (b[2] - a[2] * a[2])) * (float)N / ((float)(N - 1u) * 3.f);
  err[x] = fminf(e, 1.0f);
}
hipError_t launch_emp_error(const float* result, const float* sq, const uint32_t* sample_n, uint32_t n, float* err,
                            hipStream_t st) {
  if (!result || !sq || !sample_n || !err) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  k_emp_error<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(result, sq, sample_n, n, err);
  return hipGetLastError();
}

// Maximum with 0 as the initial value: non-negative floats order as their bits.
__global__ __launch_bounds__(kBlock) void k_error_max(const float* __restrict__ e, uint32_t n,
                                                      uint32_t* __restrict__ out) {
  uint32_t m = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    m = max(m, __float_as_uint(fmaxf(e[i], 0.f)));
  for (int o = 32; o >= 1; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}
hipError_t launch_error_max(const float* err, uint32_t n, float* max_out, hipStream_t st) {
  if (!err || !max_out) return hipErrorInvalidValue;
  k_fill_u32<<<1, 1, 0, st>>>((uint32_t*)max_out, 0u);
  uint32_t nb = (n + kBlock - 1) / kBlock;
  if (nb > 1024) nb = 1024;
  if (nb) k_error_max<<<nb, kBlock, 0, st>>>(err, n, (uint32_t*)max_out);
  return hipGetLastError();
}

__global__ void k_round_init(RoundState* rs, const float* maxe, uint32_t gen, uint32_t bisect, float desired,
                             float min_prob) {
  RoundState s;
  s.maxe = *maxe;
  s.left = 0.f;
  s.right = s.maxe * 2.f;
  s.window = bisect ? (s.left + s.right) * 0.5f : s.maxe;
  s.accum = 0;
  s.gen = gen;
  s.done = gen == 0 ? 1u : 0u;
  s.rounds = 0;
  s.bisect = bisect;
  s.desired = desired;
  s.min_prob = min_prob;
  s.pad = 0;
  *rs = s;
}
__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0u;
}
hipError_t launch_round_init(RoundState* rs, const float* maxe, uint32_t gen, uint32_t bisect, float desired,
                             float min_prob, uint32_t* counts, uint32_t n, hipStream_t st) {
  if (!rs || !maxe || !counts) return hipErrorInvalidValue;
  k_round_init<<<1, 1, 0, st>>>(rs, maxe, gen, bisect, desired, min_prob);
  if (n) k_zero_u32<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(counts, n);
  return hipGetLastError();
}

// pt_genTempAdapPathQueue_kernel: keep unless u * window > max(e, min_prob),
// u from the stream of (frame, slots filled so far) at the pixel.
__global__ __launch_bounds__(kBlock) void k_round_keep(uint64_t seed, uint32_t frame, const float* __restrict__ err,
                                                       uint32_t n, const RoundState* __restrict__ rs,
                                                       uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || rs->done) return;
  uint64_t s = 0;
  const uint64_t inc = ((uint64_t)i << 1) | 1u;
  pcg32_seed(s, inc, stream_key(seed, frame, 3u, rs->accum));
  const float u = uniform01(s, inc);
  flags[i] = (u * rs->window > fmaxf(err[i], rs->min_prob)) ? 0u : 1u;
}
// Kept pixels of rank < remaining slots take one sample each.
__global__ __launch_bounds__(kBlock) void k_round_apply(const uint32_t* __restrict__ flags,
                                                        const uint32_t* __restrict__ rank, uint32_t n,
                                                        const RoundState* __restrict__ rs,
                                                        const uint32_t* __restrict__ kept,
                                                        uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || rs->done) return;
  const uint32_t take = min(*kept, rs->gen - rs->accum);
  if (flags[i] && rank[i] < take) counts[i] += 1u;
}
// remove_if + size bookkeeping + the bisection of AdapBilatBinary.cu:1363-1371.
__global__ void k_round_update(RoundState* rs, const uint32_t* kept) {
  RoundState s = *rs;
  if (s.done) return;
  const uint32_t take = min(*kept, s.gen - s.accum);
  s.accum += take;
  s.rounds++;
  if (take == 0u) {
    s.done = 1u;
  } else if (s.bisect) {
    if ((float)take < s.desired) s.right = s.window;
    else if ((float)take > s.desired) s.left = s.window;
    s.window = (s.left + s.right) * 0.5f;
  }
  if (s.accum >= s.gen) s.done = 1u;
  *rs = s;
}
__global__ __launch_bounds__(kBlock) void k_round_fill(const RoundState* __restrict__ rs, uint32_t n,
                                                       uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t unfilled = rs->gen > rs->accum ? rs->gen - rs->accum : 0u;
  counts[i] += unfilled / n + (i < unfilled % n ? 1u : 0u);
}

// Exclusive scan of given per-element values (SRC = 1) into offsets.
__global__ __launch_bounds__(kBlock) void k_scan_local_u32(const uint32_t* __restrict__ in, uint32_t n,
                                                           uint32_t* __restrict__ offsets,
                                                           uint32_t* __restrict__ btot) {
  __shared__ uint32_t s[kBlock];
  const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
  uint32_t c[kScanPer], tsum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) {
    uint32_t i = base + k;
    c[k] = i < n ? in[i] : 0u;
    tsum += c[k];
  }
  s[threadIdx.x] = tsum;
  __syncthreads();
  for (int o = 1; o < kBlock; o <<= 1) {
    uint32_t v = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tsum;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) {
    uint32_t i = base + k;
    if (i < n) offsets[i] = run;
    run += c[k];
  }
  if (threadIdx.x == kBlock - 1) btot[blockIdx.x] = s[kBlock - 1];
}
hipError_t launch_scan_counts(const uint32_t* counts, uint32_t n, uint32_t* offsets, uint32_t* block_tot,
                              uint32_t* total, hipStream_t st) {
  if (!counts || !offsets || !block_tot || !total) return hipErrorInvalidValue;
  uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb) k_scan_local_u32<<<nb, kBlock, 0, st>>>(counts, n, offsets, block_tot);
  k_scan_totals<<<1, kBlock, 0, st>>>(block_tot, nb, total);
  if (n) k_scan_add<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(offsets, block_tot, n);
  return hipGetLastError();
}
hipError_t launch_round(const FrameArgs& fa, const float* err, uint32_t n, RoundState* rs, uint32_t* flags,
                        uint32_t* rank, uint32_t* block_tot, uint32_t* kept, uint32_t* counts, hipStream_t st) {
  if (!err || !rs || !flags || !rank || !block_tot || !kept || !counts || n == 0) return hipErrorInvalidValue;
  const uint32_t g = (n + kBlock - 1) / kBlock;
  k_round_keep<<<g, kBlock, 0, st>>>(fa.seed, fa.frame0, err, n, rs, flags);
  hipError_t e = launch_scan_counts(flags, n, rank, block_tot, kept, st);
  if (e != hipSuccess) return e;
  k_round_apply<<<g, kBlock, 0, st>>>(flags, rank, n, rs, kept, counts);
  k_round_update<<<1, 1, 0, st>>>(rs, kept);
  return hipGetLastError();
}
hipError_t launch_round_fill(const RoundState* rs, uint32_t n, uint32_t* counts, hipStream_t st) {
  if (!rs || !counts) return hipErrorInvalidValue;
  if (n) k_round_fill<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(rs, n, counts);
  return hipGetLastError();
}


// ---------------------------------------------------------------- per-pixel path caps
// A pixel that already has maxp paths gets no error weight (the allocation
// gives it nothing); counts of a round-based pass stop at maxp - paths.
__global__ __launch_bounds__(kBlock) void k_cap_error(float* __restrict__ err, const uint32_t* __restrict__ paths,
                                                      uint32_t maxp, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n && paths[i] >= maxp) err[i] = 0.f;
}
__global__ __launch_bounds__(kBlock) void k_cap_counts(uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ paths, uint32_t maxp,
                                                       uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) counts[i] = cap_count(counts[i], paths, maxp, i);
}
// paths += counts (an adaptive pass), or += 1 (a warm-up frame); reset first
// when `reset`.
__global__ __launch_bounds__(kBlock) void k_add_paths(uint32_t* __restrict__ paths,
                                                      const uint32_t* __restrict__ counts, uint32_t n, int reset) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) paths[i] = (reset ? 0u : paths[i]) + (counts ? counts[i] : 1u);
}
hipError_t launch_cap_error(float* err, const uint32_t* paths, uint32_t maxp, uint32_t n, hipStream_t st) {
  if (!maxp || n == 0) return hipSuccess;
  k_cap_error<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(err, paths, maxp, n);
  return hipGetLastError();
}
hipError_t launch_cap_counts(uint32_t* counts, const uint32_t* paths, uint32_t maxp, uint32_t n, hipStream_t st) {
  if (!maxp || n == 0) return hipSuccess;
  k_cap_counts<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(counts, paths, maxp, n);
  return hipGetLastError();
}
hipError_t launch_add_paths(uint32_t* paths, const uint32_t* counts, uint32_t n, int reset, hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_add_paths<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(paths, counts, n, reset);
  return hipGetLastError();
}
}  // namespace cpt
