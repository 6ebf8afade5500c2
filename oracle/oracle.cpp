// oracle/oracle.cpp -- CPU restatement of CudaPT's streaming-BDPT hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load liboracle.so, and only as the checker (or as the timed
// CPU baseline).  The product path fails loudly when its HIP library is absent.
//
// What each function restates (paths relative to the reference checkout):
//   pcg32_*            PCG32, O'Neill 2014 (pcg-c-basic), replaces cuRAND XORWOW
//                      (cudaRTBDPTStream.cu:552,633) -- pinned by the published
//                      pcg32-demo known-answer vector (tests/golden).
//   splitmix64         Steele/Lea/Flood 2014 finaliser -- pinned by its KAT.
//   bvh_build_literal  NPBVHHelper::CreateBVH / InitialBVHNode /
//                      InitialCompactBVH (src/bvhhelper.cpp:33-311), kept in
//                      the reference's per-bin loop form -- pinned by
//                      hand-worked 2..8 triangle cases and structural invariants.
//   intersect_tri      CURay::IntersectTri (src/cudaRTCommon.cu:42-64) with the
//                      device arithmetic of include/cudahelper.h:127-198 --
//                      pinned by closed-form triangles.
//   intersect_aabb     CURay::IntersectAABB (src/cudaRTCommon.cu:24-40).
//   trace_refbin       TracePrimitive (src/cudaRTCommon.cu:66-127): right-first
//                      DFS, 32-entry stack dropping pushes at 29, 2048-pop cap,
//                      strict '<' first-found minimum.
//   trace_brute        All-triangle minimum of (t, original triangle id): the
//                      tie-defined closest hit the product promises.
//   trace_defined      Same answer as trace_brute through the binary BVH with
//                      conservative box culling -- pinned against trace_brute.
//   render_frames      cudaRTBDPTStream::Render (src/cudaRTBDPTStream.cu:
//                      171-951): light-vertex cache every 3rd frame, eye bounce,
//                      light-vertex connection (N += 4), running-mean apply.
//                      Pinned by the emitter-only and all-black scenes and the
//                      hand-checked single-connection case; the FP32 GPU run is
//                      held to relative MSE 1e-4 of the FP64 run of this code.
//   d_ggx, vis_smith,  D_GGX, Vis_SmithJointApprox, F_Schlick, ImportanceSampleGGX
//   f_schlick,         (src/cudaRTCommon.cu:284-345) and GetShadingResult
//   sample_ggx,        (src/cudaRTBDPTStream.cu:642-663) -- pinned by GGX
//   shading            normalisation quadrature, the normal-incidence value
//                      1/(pi r^4), a KS test of the sampler against the
//                      quadrature CDF of D (n.h), Schlick and Smith hand values,
//                      a Lambert white furnace and the specular lobe's missing
//                      1/(4 NoL NoV) by hand (tests/test_oracle_shading.py).
//   sample_bounce      the eye/light bounce's BRDF sampling
//                      (src/cudaRTBDPTStream.cu:395-502) -- pinned by a KS test
//                      of the r2sin = 1 - r2cos^2 hemisphere against its closed
//                      CDF, branch frequencies lerp(|F|, 1, metallic) and
//                      transparency, and each branch's weight in closed form.
//   eye_path RR        Russian roulette u > |T| |f| without reweighting
//                      (:515-520) -- pinned by the mean path length in a closed
//                      grey box.
//   light_pass         (:180-363,545-617) -- pinned by a hand-worked two-bounce
//                      subpath: cache slot order, irrad zeroed before a
//                      non-diffuse bounce (:331-332), the sample update.
//   bilateral_posnorm  cudaBilateralFilter::updateGaussian and
//                      bilaterial_posnorm(emit)_kernel
//                      (src/cudaBilateralFilter.cu:5-142), including the
//                      mixed int/uint clamp that sends low-side taps to the
//                      high edge -- pinned by radius-0 identity, constant-image
//                      fixed point and the large-sigma Gaussian closed form.
//   square_error       pt_calculateSquareError_kernel
//                      (src/cudaRTBDPTStreamAdapBilatAtomic.cu:1143-1156).
//   empirical_error    pt_calculateEmpiricalSquareError_kernel
//                      (src/cudaRTBDPTStreamAdapEmp.cu:949-960).
//   alloc_proportional pt_genAdapPathQueue_kernel sample counts
//                      (src/cudaRTBDPTStreamAdapBilatAtomic.cu:817-842).
//
// Precision: every numeric routine is a template over R in {float, double}.
// R = float reproduces the device operation sequence (explicit fma nesting in
// dot/cross, correctly rounded divide, round-down sqrt in length()); R = double
// is the FP64 run the FP32 GPU radiance is judged against.  Random uniforms are
// 24-bit ((x >> 8) + 1) * 2^-24, exact in both precisions.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).

#include <cmath>
#include <type_traits>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint32_t u32;
typedef int32_t i32;
typedef uint64_t u64;

namespace orc {

// ---------------------------------------------------------------- constants
// include/cudahelper.h:27-30
static const double kEpsilon = 1e-9;     // M_EPSILON
static const double kFltEps = 1e-5;      // M_FLT_EPSILON
static const double kBiasEps = 1e-4;     // M_FLT_BIAS_EPSILON
static const double kInf = 1e20;         // M_INF
static const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- RNG
struct Pcg32 { u64 state, inc; };
static inline u32 pcg32_next(Pcg32& r) {
  u64 old = r.state;
  r.state = old * 6364136223846793005ULL + r.inc;
  u32 xs = (u32)(((old >> 18u) ^ old) >> 27u);
  u32 rot = (u32)(old >> 59u);
  return (xs >> rot) | (xs << ((-rot) & 31u));
}
static inline void pcg32_seed(Pcg32& r, u64 initstate, u64 initseq) {
  r.state = 0u;
  r.inc = (initseq << 1u) | 1u;
  pcg32_next(r);
  r.state += initstate;
  pcg32_next(r);
}
static inline u64 splitmix64_mix(u64 z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// Stream key for (seed, frame, domain, sample): the counter-based analogue of
// curand_init(WangHash(frameN) + index, 0, 0) (cudaRTBDPTStream.cu:552,633).
static inline u64 stream_key(u64 seed, u32 frame, u32 domain, u32 sample) {
  return splitmix64_mix(seed + 0x9E3779B97F4A7C15ULL * (u64)(frame + 1) +
                        ((u64)sample << 32) + ((u64)domain << 60));
}
// Uniform in (0, 1]: same support as curand_uniform, exact in FP32 and FP64.
template <class R> static inline R uniform(Pcg32& r) {
  return (R)((pcg32_next(r) >> 8) + 1u) * (R)(1.0 / 16777216.0);
}
enum { DOMAIN_EYE = 1, DOMAIN_LIGHT = 2 };

// ---------------------------------------------------------------- vectors
template <class R> struct V3 { R x, y, z; };
template <class R> static inline V3<R> mk(R x, R y, R z) { V3<R> v; v.x = x; v.y = y; v.z = z; return v; }
template <class R> static inline R fm(R a, R b, R c) { return std::fma(a, b, c); }
template <class R> static inline V3<R> operator+(V3<R> a, V3<R> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> static inline V3<R> operator-(V3<R> a, V3<R> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> static inline V3<R> operator*(V3<R> a, R s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class R> static inline V3<R> operator*(R s, V3<R> a) { return a * s; }
// A vector over a scalar.  FP32 runs use the device's form: one correctly
// rounded reciprocal, then three products (within an ulp of the reference's
// three divisions); FP64 runs divide.
template <class R> static inline V3<R> operator/(V3<R> a, R s) {
  if constexpr (std::is_same<R, float>::value) {
    const float r = 1.f / s;
    return mk(a.x * r, a.y * r, a.z * r);
  } else {
    return mk(a.x / s, a.y / s, a.z / s);
  }
}
// ---------------------------------------------------------------- trigonometry
// FP32 runs evaluate sin/cos, atan2 and acos with the device's operation
// sequence (devmath.h): Cephes' (S. L. Moshier) sinf/cosf octant reduction
// and minimax polynomials for x >= 0, Cephes' atanf core on [0, 1] with the
// octants of atan2, and acos(y) = atan2(sqrt((1 - y)(1 + y)), y).  FP64 runs
// use libm.
static inline void sincos_f(float x, float& s, float& c) {
  int j = (int)(x * 1.27323954473516f);
  j += j & 1;
  const float y = (float)j;
  float r = std::fma(-y, 0.78515625f, x);
  r = std::fma(-y, 2.4187564849853515625e-4f, r);
  r = std::fma(-y, 3.77489497744594108e-8f, r);
  const float z = r * r;
  const float ps = std::fma(std::fma(std::fma(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f), z * r, r);
  const float pc = std::fma(std::fma(std::fma(2.443315711809948e-5f, z, -1.388731625493765e-3f), z,
                                     4.166664568298827e-2f),
                            z * z, std::fma(-0.5f, z, 1.f));
  const int q = (j >> 1) & 3;
  s = (q & 1) ? pc : ps;
  c = (q & 1) ? ps : pc;
  if (q >= 2) s = -s;
  if (q == 1 || q == 2) c = -c;
}
static inline float atan01_f(float t) {
  float y0 = 0.f;
  if (t > 0.4142135623730950f) {
    y0 = 0.78539816339744831f;
    t = (t - 1.f) / (t + 1.f);
  }
  const float z = t * t;
  const float p = std::fma(std::fma(std::fma(8.05374449538e-2f, z, -1.38776856032e-1f), z, 1.99777106478e-1f), z,
                           -3.33329491539e-1f);
  return y0 + std::fma(p * z, t, t);
}
static inline float atan2_f(float y, float x) {
  const float ax = std::fabs(x), ay = std::fabs(y);
  float a;
  if (ay <= ax) a = ax > 0.f ? atan01_f(ay / ax) : 0.f;
  else a = 1.57079632679489662f - atan01_f(ax / ay);
  if (x < 0.f) a = 3.14159265358979324f - a;
  return y < 0.f ? -a : a;
}
template <class R> static inline void sincos_r(R x, R& s, R& c) {
  if constexpr (std::is_same<R, float>::value) sincos_f(x, s, c);
  else { s = std::sin(x); c = std::cos(x); }
}
template <class R> static inline R atan2_r(R y, R x) {
  if constexpr (std::is_same<R, float>::value) return atan2_f(y, x);
  else return std::atan2(y, x);
}
template <class R> static inline R acos_r(R y) {
  if constexpr (std::is_same<R, float>::value) return atan2_f(std::sqrt((1.f - y) * (1.f + y)), y);
  else return std::acos(y);
}

template <class R> static inline V3<R> vmul(V3<R> a, V3<R> b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> static inline V3<R> vdiv(V3<R> a, V3<R> b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> static inline V3<R> vmax(V3<R> a, V3<R> b) { return mk(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
template <class R> static inline V3<R> vmin(V3<R> a, V3<R> b) { return mk(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
// cudahelper.h:179-198 -- explicit fma nesting of the device dot/cross.
template <class R> static inline R dot(V3<R> a, V3<R> b) { return fm(a.x, b.x, fm(a.y, b.y, a.z * b.z)); }
template <class R> static inline V3<R> cross(V3<R> a, V3<R> b) {
  return mk(fm(a.y, b.z, -(a.z * b.y)), fm(a.z, b.x, -(a.x * b.z)), fm(a.x, b.y, -(a.y * b.x)));
}
// cudahelper.h:292-299: __fsqrt_rd of the (contracted) sum of squares.
template <class R> static inline R sqrt_rd(R x) {
  R s = std::sqrt(x);
  if (s > (R)0 && std::fma(s, s, -x) > (R)0) s = std::nextafter(s, (R)0);
  return s;
}
template <class R> static inline R length(V3<R> a) { return sqrt_rd(fm(a.z, a.z, fm(a.y, a.y, a.x * a.x))); }
template <class R> static inline V3<R> normalize(V3<R> a) { return a / length(a); }
template <class R> static inline R lerp(R x, R y, R d) { return x + (y - x) * d; }
template <class R> static inline V3<R> vlerp(V3<R> x, V3<R> y, R d) { return x + (y - x) * d; }
template <class R> static inline R saturate(R x) { return x < (R)0 ? (R)0 : (x > (R)1 ? (R)1 : x); }
template <class R> static inline V3<R> cvt(const float* p) { return mk((R)p[0], (R)p[1], (R)p[2]); }
template <class R> static inline bool isinf3(V3<R> a) { return std::isinf(a.x) || std::isinf(a.y) || std::isinf(a.z); }

// smallest float >= 1e-9: for a float x, x < kDetEpsF  <=>  (double)x < 1e-9.
static float det_eps_f() {
  float f = (float)kEpsilon;
  if ((double)f < kEpsilon) f = std::nextafter(f, 1.0f);
  return f;
}
template <class R> static inline R det_eps();
template <> inline float det_eps<float>() { static const float v = det_eps_f(); return v; }
template <> inline double det_eps<double>() { return kEpsilon; }

// ---------------------------------------------------------------- scene
struct Material {            // RTMaterial field order (include/raytracer.h:82-100)
  float diffuse[3];
  float emissive[3];
  i32 diffuseTexId, normalTexId, emissiveTexId, matType;
  float transparency, subsurface, specular, metallic, roughness, anisotropic;
  float sheen, sheenTint, clearcoat, clearcoatGloss;
};
static_assert(sizeof(Material) == 80, "RTMaterial is 80 bytes");

struct Scene {
  u32 nv = 0, nt = 0;
  std::vector<float> pos, nrm;       // 3 per vertex
  std::vector<u32> tri;              // 3 per triangle, BVH leaf order
  std::vector<u32> tmat;             // material per triangle, leaf order
  std::vector<u32> orig;             // leaf order -> original triangle id
  std::vector<Material> mats;
  std::vector<float> p0, e0, e1;     // 3 per leaf-order triangle (float, host-subtracted)
  // compact binary BVH (bvhhelper.h:45-54)
  u32 nodeN = 0;
  std::vector<float> bmin, bmax;     // 3 per node
  std::vector<u32> offOrTSTN;        // 2 per node
  std::vector<u32> lightTri;         // leaf-order emissive triangles
  std::vector<float> uv;             // 2 per vertex (zeros unless set)
  std::vector<float> tex;            // RTScene::m_pTextures: RGBA texels, concatenated
  std::vector<u32> texDesc;          // 3 per texture: first texel, width, height
  std::vector<float> env;            // environment light: envW x envH RGB texels (empty: none)
  u32 envW = 0, envH = 0;
};

// ---------------------------------------------------------------- BVH build
// src/bvhhelper.cpp, restated literally (float host arithmetic, no contraction).
struct F3 { float x, y, z; float& operator[](int i) { return (&x)[i]; } float operator[](int i) const { return (&x)[i]; } };
static inline F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline F3 f3min(F3 a, F3 b) { return f3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
static inline F3 f3max(F3 a, F3 b) { return f3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
struct Box { F3 mn, mx; };
static inline Box empty_box() { Box b; b.mn = f3((float)kInf, (float)kInf, (float)kInf); b.mx = f3((float)-kInf, (float)-kInf, (float)-kInf); return b; }
static inline Box merge(Box a, Box b) { Box r; r.mn = f3min(a.mn, b.mn); r.mx = f3max(a.mx, b.mx); return r; }
static inline Box merge(Box a, F3 p) { Box r; r.mn = f3min(a.mn, p); r.mx = f3max(a.mx, p); return r; }
// bvhhelper.cpp:33-39
static float surface_area(const Box& b) {
  F3 d = f3(b.mx.x - b.mn.x, b.mx.y - b.mn.y, b.mx.z - b.mn.z);
  float len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
  if (d.x < 0 || d.y < 0 || d.z < 0 || (double)len < kEpsilon) return 0;
  return d.x * d.y + d.x * d.z + d.y * d.z;
}
struct BNode { BNode* c[2] = {nullptr, nullptr}; Box bound = empty_box(); u32 triStart = 0, triN = 0, desN = 0; };
struct BTri { F3 centroid; Box bound; };
static const int kBins = 16;

static void free_tree(BNode* n) { if (!n) return; free_tree(n->c[0]); free_tree(n->c[1]); delete n; }

// bvhhelper.cpp:55-237
static void build_node(BNode* node, const BTri* tris, u32* order, u32 triStart, u32 triN, u32 depthBudget) {
  if (triN <= 0) return;
  Box cb = empty_box();
  for (u32 i = 0; i < triN; i++) {
    u32 ind = order[triStart + i];
    node->bound = merge(node->bound, tris[ind].bound);
    cb = merge(cb, tris[ind].centroid);
  }
  F3 dia = f3(cb.mx.x - cb.mn.x, cb.mx.y - cb.mn.y, cb.mx.z - cb.mn.z);
  if (dia.x < 0 || dia.y < 0 || dia.z < 0) return;
  float dlen = std::sqrt(dia.x * dia.x + dia.y * dia.y + dia.z * dia.z);
  if (triN == 1 || depthBudget <= 1 || (double)dlen < kEpsilon) {
    node->triStart = triStart; node->triN = triN; node->desN = 0; return;
  }
  u32 axis = 0;
  if (std::fabs(dia.y) > std::fabs(dia.x)) axis = 1;
  if (std::fabs(dia.z) > std::fabs(dia.y) && std::fabs(dia.z) > std::fabs(dia.x)) axis = 2;
  float len = std::fabs(dia[axis]);
  // bins, literal per-bin scan (bvhhelper.cpp:86-111)
  u32 binN[kBins]; Box binB[kBins];
  for (int i = 0; i < kBins; i++) {
    binN[i] = 0; binB[i] = empty_box();
    float s = cb.mn[axis] + (float)i * len / (float)kBins;
    float e = cb.mn[axis] + (float)(i + 1) * len / (float)kBins;
    if (i + 1 == kBins) e = (float)((double)cb.mx[axis] + kBiasEps);
    for (u32 j = 0; j < triN; j++) {
      u32 ind = order[triStart + j];
      float p = tris[ind].centroid[axis];
      if (p >= s && p < e) { binN[i]++; binB[i] = merge(binB[i], tris[ind].centroid); }
    }
  }
  // costs (bvhhelper.cpp:120-164): heu[k] = left bins 0..k, right bins k+1..15
  float nodeSA = surface_area(node->bound);
  float heu[kBins];
  for (int pass = 0; pass < 2; pass++) {
    u32 start = pass == 0 ? 0 : kBins, end = kBins / 2, step = pass == 0 ? 1u : (u32)-1;
    u32 mainN = 0; Box mainB = empty_box();
    for (u32 j = start; j != end; j += step) {
      if (j < (u32)kBins) { mainN += binN[j]; mainB = merge(mainB, binB[j]); }
      u32 opN = 0; Box opB = empty_box();
      for (u32 k = j + step; k < (u32)kBins; k += step) { opN += binN[k]; opB = merge(opB, binB[k]); }
      float c = 0.125f + ((float)mainN * surface_area(mainB) + (float)opN * surface_area(opB)) / nodeSA;
      if (pass == 0) heu[j] = c; else heu[j - 1] = c;
    }
  }
  float minCost = (float)kInf, minBin = (float)(kBins - 1);
  for (int i = 0; i < kBins; i++) if (heu[i] < minCost) { minCost = heu[i]; minBin = (float)i; }
  // partition (bvhhelper.cpp:190-211)
  float split = cb.mn[axis] + (minBin + 1) * len / (float)kBins;
  u32 left = triStart, right = triStart + triN - 1, mid;
  while (left < right) {
    while (tris[order[left]].centroid[axis] <= split && left < right) left++;
    while (tris[order[right]].centroid[axis] > split && left < right) right--;
    if (left < right) { std::swap(order[left], order[right]); left++; right--; }
  }
  mid = (tris[order[left]].centroid[axis] <= split) ? left : left - 1;
  BNode* l = new BNode(); BNode* r = new BNode();
  node->c[0] = l; node->c[1] = r;
  build_node(l, tris, order, triStart, mid - triStart + 1, depthBudget - 1);
  build_node(r, tris, order, mid + 1, triN - (mid - triStart + 1), depthBudget - 1);
  node->desN = 2 + l->desN + r->desN;
}

// bvhhelper.cpp:285-298
static u32 compact_node(const BNode* n, Scene& s, u32 cur) {
  s.bmin[cur * 3 + 0] = n->bound.mn.x; s.bmin[cur * 3 + 1] = n->bound.mn.y; s.bmin[cur * 3 + 2] = n->bound.mn.z;
  s.bmax[cur * 3 + 0] = n->bound.mx.x; s.bmax[cur * 3 + 1] = n->bound.mx.y; s.bmax[cur * 3 + 2] = n->bound.mx.z;
  if (!n->c[0] || !n->c[1]) { s.offOrTSTN[cur * 2] = n->triStart; s.offOrTSTN[cur * 2 + 1] = n->triN; return 1; }
  s.offOrTSTN[cur * 2 + 1] = 0;
  u32 first = compact_node(n->c[0], s, cur + 1);
  s.offOrTSTN[cur * 2] = first + 1;
  return compact_node(n->c[1], s, cur + first + 1) + first + 1;
}

// bvhhelper.cpp:239-283 + raytracer.cpp:255-281 (reorder, intersection data)
static bool build_scene_bvh(Scene& s, const float* pos, const u32* idx, u32 maxDepth, std::vector<u32>& order) {
  u32 nt = s.nt;
  std::vector<BTri> tris(nt);
  order.resize(nt);
  for (u32 i = 0; i < nt; i++) {
    order[i] = i;
    F3 a = f3(pos[idx[i * 3] * 3], pos[idx[i * 3] * 3 + 1], pos[idx[i * 3] * 3 + 2]);
    F3 b = f3(pos[idx[i * 3 + 1] * 3], pos[idx[i * 3 + 1] * 3 + 1], pos[idx[i * 3 + 1] * 3 + 2]);
    F3 c = f3(pos[idx[i * 3 + 2] * 3], pos[idx[i * 3 + 2] * 3 + 1], pos[idx[i * 3 + 2] * 3 + 2]);
    tris[i].centroid = f3(((a.x + b.x) + c.x) / 3.f, ((a.y + b.y) + c.y) / 3.f, ((a.z + b.z) + c.z) / 3.f);
    tris[i].bound.mn = f3min(c, f3min(b, a));
    tris[i].bound.mx = f3max(c, f3max(b, a));
  }
  if (nt == 0) return false;
  BNode* root = new BNode();
  build_node(root, tris.data(), order.data(), 0, nt, maxDepth);
  s.nodeN = root->desN + 1;
  s.bmin.assign(s.nodeN * 3, 0.f); s.bmax.assign(s.nodeN * 3, 0.f); s.offOrTSTN.assign(s.nodeN * 2, 0u);
  bool ok = compact_node(root, s, 0) == s.nodeN;
  free_tree(root);
  return ok;
}

// ---------------------------------------------------------------- intersection
// cudaRTCommon.cu:42-64 (IntersectTri) with the device arithmetic.
template <class R>
static inline R intersect_tri(V3<R> o, V3<R> d, V3<R> p0, V3<R> e0, V3<R> e1, R eps, bool cull, R& u, R& v) {
  if (cull && dot(cross(e0, e1), d) > (R)0) return (R)kInf;
  V3<R> de2 = cross(d, e1);
  R div = dot(de2, e0);
  if (std::fabs(div) < det_eps<R>()) return (R)kInf;
  div = (R)1 / div;
  V3<R> t = (o + d * eps) - p0;
  V3<R> te1 = cross(t, e0);
  R rT = dot(te1, e1) * div;
  if (rT < (R)0) return (R)kInf;
  u = dot(de2, t) * div;
  if (u < (R)0 || u > (R)1) return (R)kInf;
  v = dot(te1, d) * div;
  if (v < (R)0 || (u + v) > (R)1) return (R)kInf;
  return rT;
}
// cudahelper.h:119-125
template <class R> static inline R escape_zero(R v, R e) { if (std::fabs(v) < e) v = (v > (R)0) ? v + e : v - e; return v; }
// cudaRTCommon.cu:24-40
template <class R>
static inline R intersect_aabb(V3<R> o, V3<R> d, V3<R> mn, V3<R> mx) {
  R e = (R)(float)kEpsilon;
  V3<R> rd = mk((R)1 / escape_zero(d.x, e), (R)1 / escape_zero(d.y, e), (R)1 / escape_zero(d.z, e));
  V3<R> t0 = vmul(mn - o, rd), t1 = vmul(mx - o, rd);
  V3<R> lo = vmin(t0, t1), hi = vmax(t0, t1);
  R minmax = std::fmin(std::fmin(hi.x, hi.y), hi.z);
  R maxmin = std::fmax(std::fmax(lo.x, lo.y), lo.z);
  if (minmax >= maxmin && (minmax > (R)0 || maxmin > (R)0)) return ((double)maxmin > kEpsilon) ? maxmin : (R)0;
  return (R)kInf;
}

struct Hit { u32 tri; float t, u, v; };   // tri = original id, 0xFFFFFFFF on miss

// Conservative slab test for the defined traversal: exact-form slab distances
// with the far distance widened by 1 + 2*gamma(3) (Ize, "Robust BVH Ray
// Traversal", JCGT 2013), so a box that the ray touches is never culled.
template <class R>
static inline R conservative_aabb(V3<R> o, V3<R> d, V3<R> mn, V3<R> mx) {
  R e = (R)(float)kEpsilon;
  V3<R> rd = mk((R)1 / escape_zero(d.x, e), (R)1 / escape_zero(d.y, e), (R)1 / escape_zero(d.z, e));
  V3<R> t0 = vmul(mn - o, rd), t1 = vmul(mx - o, rd);
  V3<R> lo = vmin(t0, t1), hi = vmax(t0, t1);
  R tnear = std::fmax(std::fmax(lo.x, lo.y), std::fmax(lo.z, (R)0));
  R tfar = std::fmin(std::fmin(hi.x, hi.y), hi.z) * (R)1.0000004;
  return tnear <= tfar ? tnear : std::numeric_limits<R>::infinity();
}

template <class R> struct HitR { u32 leaf; u32 orig; R t, u, v; };

// Tie-defined acceptance: strictly closer, or equal t with a smaller original id.
template <class R> static inline bool better(R t, u32 id, const HitR<R>& h) {
  if (t < h.t) return true;
  return t == h.t && h.leaf != 0xFFFFFFFFu && id < h.orig;
}

// Conservative box slack: covers the eps origin offset of IntersectTri and the
// rounding of the slab test, so culling never drops the tie-defined minimum.
template <class R> static inline R cull_slack(R best, R eps) { return eps + best * (R)1e-4 + (R)1e-5; }

template <class R>
static bool trace_brute(const Scene& s, V3<R> o, V3<R> d, R maxDist, R eps, bool cull, HitR<R>& h) {
  h.leaf = 0xFFFFFFFFu; h.orig = 0xFFFFFFFFu; h.t = maxDist; h.u = h.v = 0;
  for (u32 i = 0; i < s.nt; i++) {
    R u = 0, v = 0;
    R t = intersect_tri(o, d, cvt<R>(&s.p0[i * 3]), cvt<R>(&s.e0[i * 3]), cvt<R>(&s.e1[i * 3]), eps, cull, u, v);
    if (t >= (R)0 && better(t, s.orig[i], h)) { h.leaf = i; h.orig = s.orig[i]; h.t = t; h.u = u; h.v = v; }
  }
  return h.leaf != 0xFFFFFFFFu;
}

template <class R>
static bool trace_defined(const Scene& s, V3<R> o, V3<R> d, R maxDist, R eps, bool cull, HitR<R>& h, bool any_hit = false) {
  h.leaf = 0xFFFFFFFFu; h.orig = 0xFFFFFFFFu; h.t = maxDist; h.u = h.v = 0;
  if (s.nodeN == 0) return false;
  u32 stack[128]; int sp = 0; stack[0] = 0;
  while (sp >= 0) {
    u32 n = stack[sp--];
    R tb = conservative_aabb(o, d, cvt<R>(&s.bmin[n * 3]), cvt<R>(&s.bmax[n * 3]));
    if (!(tb <= h.t + cull_slack(h.t, eps))) continue;
    u32 a = s.offOrTSTN[n * 2], cnt = s.offOrTSTN[n * 2 + 1];
    if (cnt == 0) {
      if (s.bmin[n * 3] > s.bmax[n * 3]) continue;       // empty leaf of a degenerate split
      stack[++sp] = n + 1; stack[++sp] = n + a;
    } else {
      for (u32 i = a; i < a + cnt; i++) {
        R u = 0, v = 0;
        R t = intersect_tri(o, d, cvt<R>(&s.p0[i * 3]), cvt<R>(&s.e0[i * 3]), cvt<R>(&s.e1[i * 3]), eps, cull, u, v);
        if (t >= (R)0 && better(t, s.orig[i], h)) {
          h.leaf = i; h.orig = s.orig[i]; h.t = t; h.u = u; h.v = v;
          if (any_hit) return true;
        }
      }
    }
  }
  return h.leaf != 0xFFFFFFFFu;
}

// cudaRTCommon.cu:66-127, literal (reference tie and cap behaviour).
struct RefStats { u32 stack_drops, cap_hits; };
template <class R>
static bool trace_refbin(const Scene& s, V3<R> o, V3<R> d, R maxDist, R eps, bool cull, HitR<R>& h, RefStats& st) {
  R best = maxDist; u32 id = 0; R hu = 0, hv = 0;
  u32 stack[32]; i32 sp = 0; stack[0] = 0; u32 traced = 0;
  while (sp >= 0 && traced++ < 2048u) {
    u32 n = stack[sp--];
    R tb = intersect_aabb(o, d, cvt<R>(&s.bmin[n * 3]), cvt<R>(&s.bmax[n * 3]));
    if (tb >= (R)0 && tb < best) {
      u32 a = s.offOrTSTN[n * 2], cnt = s.offOrTSTN[n * 2 + 1];
      if (cnt == 0) {
        if (sp < 32 - 3) { stack[++sp] = n + 1; stack[++sp] = n + a; }
        else st.stack_drops++;
      } else {
        for (u32 i = a; i < a + cnt; i++) {
          R u = 0, v = 0;
          R t = intersect_tri(o, d, cvt<R>(&s.p0[i * 3]), cvt<R>(&s.e0[i * 3]), cvt<R>(&s.e1[i * 3]), eps, cull, u, v);
          if (t >= (R)0 && t < best) { best = t; id = i; hu = u; hv = v; }
        }
      }
    }
  }
  if (sp >= 0) st.cap_hits++;
  if (best < maxDist) { h.leaf = id; h.orig = s.orig[id]; h.t = best; h.u = hu; h.v = hv; return true; }
  h.leaf = 0xFFFFFFFFu; h.orig = 0xFFFFFFFFu; h.t = maxDist; h.u = h.v = 0;
  return false;
}

// TraceDepth (cudaRTCommon.cu:184-231) as written, over the reference tree:
// right child popped first, a 32-entry command stack, at most 2048 pops.
// depth = stack pointer + 2 at the last leaf box entered, and at least stack
// pointer + 1 at every box missed; returns depth != 0.
template <class R>
static bool trace_depth_literal(const Scene& s, V3<R> o, V3<R> d, R maxDist, u32& depth, bool& leaf) {
  R minI = maxDist;
  u32 cmd[32]; cmd[0] = 0;
  i32 sp = 0;
  depth = 0; leaf = false;
  u32 traced = 0;
  if (s.nodeN == 0) return false;
  while (sp >= 0 && traced++ < 2048u) {
    u32 n = cmd[sp--];
    R tb = intersect_aabb(o, d, cvt<R>(&s.bmin[n * 3]), cvt<R>(&s.bmax[n * 3]));
    if (tb >= (R)0 && tb < minI) {
      u32 a = s.offOrTSTN[n * 2], cnt = s.offOrTSTN[n * 2 + 1];
      if (cnt == 0) {
        if (sp < 32 - 2) { cmd[++sp] = n + 1; cmd[++sp] = n + a; }
      } else {
        leaf = true; minI = tb; depth = (u32)(sp + 2);
      }
    } else {
      depth = std::max(depth, (u32)(sp + 1));
    }
  }
  return depth != 0;
}

// TraceDepthParent (cudaRTCommon.cu:129-182) as written: the first box at tree
// depth `spec` the right-first walk enters, and its parent; depth = -1 when
// the walk enters boxes but none at that depth; returns whether it entered any.
template <class R>
static bool trace_depth_parent_literal(const Scene& s, V3<R> o, V3<R> d, R maxDist, u32 spec, i32& depth,
                                       u32& parent) {
  R minI = maxDist;
  u32 cmd[32], par[32], dep[32];
  cmd[0] = 0; par[0] = 0; dep[0] = 0;
  i32 sp = 0;
  depth = -1; parent = 0;
  bool any = false;
  u32 traced = 0;
  if (s.nodeN == 0) return false;
  while (sp >= 0 && traced++ < 2048u) {
    i32 cp = sp--;
    u32 n = cmd[cp];
    R tb = intersect_aabb(o, d, cvt<R>(&s.bmin[n * 3]), cvt<R>(&s.bmax[n * 3]));
    if (tb >= (R)0 && tb < minI) {
      any = true;
      u32 cd = dep[cp];
      if (spec == cd) { minI = tb; depth = (i32)cd; parent = par[cp]; break; }
      u32 a = s.offOrTSTN[n * 2], cnt = s.offOrTSTN[n * 2 + 1];
      if (cnt == 0 && sp < 32 - 2) {
        ++sp; cmd[sp] = n + 1; par[sp] = n; dep[sp] = cd + 1;
        ++sp; cmd[sp] = n + a; par[sp] = n; dep[sp] = cd + 1;
      }
    }
  }
  if (!any) { depth = -1; parent = 0; }
  return any;
}

// ---------------------------------------------------------------- shading
// cudaRTCommon.cu:284-365
template <class R> static inline bool prob_rand(Pcg32& rng, R prob) {
  if ((double)prob < kBiasEps) return false;
  if ((double)prob + kBiasEps > 1.0) return true;
  return uniform<R>(rng) < prob;
}
template <class R> static inline R vis_smith(R rough, R nov, R nol) {
  R k = (rough * rough) / (R)2;
  return (nov / (nov * ((R)1 - k) + k)) * (nol / (nol * ((R)1 - k) + k));
}
template <class R> static inline R d_ggx(R rough, R noh) {
  R m = rough * rough, m2 = m * m;
  R d = (noh * m2 - noh) * noh + (R)1;
  return (R)((double)m2 / (kPi * (double)d * (double)d));
}
// x^5 of Schlick's term.  FP32 runs evaluate it as the device does: the
// power in double (x^2 exact, x^4 and x^5 rounded at 2^-53) rounded once to
// float, the correctly rounded x^5 except within ~2^-51 of a rounding
// boundary; the reference's powf is within an ulp of the same value.
template <class R> static inline R pow5(R x) {
  if constexpr (std::is_same<R, float>::value) {
    const double d = (double)x, d2 = d * d;
    return (float)(d2 * d2 * d);
  } else {
    return std::pow(x, (R)5);
  }
}
template <class R> static inline V3<R> f_schlick(V3<R> sc, R voh) {
  R fc = pow5((R)1 - voh);
  R first = (R)saturate(50.0 * (double)sc.z) * fc;
  return mk(first, first, first) + sc * ((R)1 - fc);
}
template <class R> static inline void basis(V3<R> w, V3<R>& u, V3<R>& v) {
  V3<R> a = (double)std::fabs(w.x) > 0.1 ? mk((R)0, (R)1, (R)0) : mk((R)1, (R)0, (R)0);
  u = normalize(cross(a, w));
  v = cross(w, u);
  u = cross(v, w);
}
template <class R> static inline V3<R> sample_ggx(R xi0, R xi1, R rough, V3<R> n) {
  R a = rough * rough;
  R phi = (R)(2.0 * kPi * (double)xi0);
  R ct = std::sqrt(((R)1 - xi1) / ((R)1 + (a * a - (R)1) * xi1));
  R st = std::sqrt((R)1 - ct * ct);
  R sp, cp;
  sincos_r(phi, sp, cp);
  V3<R> h = mk(st * cp, st * sp, ct);
  V3<R> u, v; basis(n, u, v);
  return (u * h.x + v * h.y) + n * h.z;
}
template <class R> static inline V3<R> lambert(V3<R> diff) { return diff * (R)(1.0 / kPi); }
// cudaRTBDPTStream.cu:642-663
template <class R>
static inline V3<R> shading(V3<R> outDir, V3<R> inDir, V3<R> irrad, V3<R> n, V3<R> diff, R metallic, R rough, R spec, R dw, R sw) {
  if (dot(n, inDir) <= (R)0) return mk((R)0, (R)0, (R)0);
  V3<R> h = normalize(outDir + inDir);
  R voh = dot(outDir, h), nov = dot(n, outDir), noh = dot(n, h), nol = dot(n, inDir);
  V3<R> f0 = vlerp(mk((R)1, (R)1, (R)1) * ((R)0.08 * spec), diff, metallic);
  V3<R> F = f_schlick(f0, voh);
  R dd = d_ggx(rough, noh), vv = vis_smith(rough, nov, nol);
  V3<R> specI = F * (dd * vv);
  V3<R> diffI = vmul(mk((R)1, (R)1, (R)1) - F, lambert(diff));
  return vmul(irrad * nol, specI * sw + diffI * dw);
}

// ---------------------------------------------------------------- Disney BRDF (extension)
// No reference counterpart: CudaPT reads the Disney parameters into RTMaterial
// (include/raytracer.h:82-100) and never uses them (src/cudaRTCommon.cu:398-405).
// The principled BRDF of Burley, "Physically Based Shading at Disney" (SIGGRAPH
// 2012 course notes, and its reference shader brdf.txt), isotropic, specular tint
// 0 (RTMaterial has none), opaque:
//   f = (1 - metallic) [ lerp(Fd, ss, subsurface) base / pi + Fsheen ]
//       + Ds Fs Gs + clearcoat / 4 * Dr Fr Gr
//   Fd = (1 + (Fd90 - 1) FL)(1 + (Fd90 - 1) FV), Fd90 = 1/2 + 2 LH^2 r
//   ss = 1.25 (Fss (1 / (NL + NV) - 1/2) + 1/2), Fss from Fss90 = LH^2 r
//   Fsheen = (1 - LH)^5 sheen lerp(1, tint, sheenTint), tint = base / lum(base)
//   Ds = GTR2(NH, a), a = max(1e-3, r^2); Fs = lerp(Cspec0, 1, (1 - LH)^5),
//   Cspec0 = lerp(0.08 specular, base, metallic); Gs = sG(NL, a) sG(NV, a),
//   sG(n, a) = 1 / (n + sqrt(a^2 + n^2 - a^2 n^2)) (Smith GGX with the 1 / 4 NL NV)
//   Dr = GTR1(NH, lerp(0.1, 0.001, clearcoatGloss)), Fr = lerp(0.04, 1, (1 - LH)^5),
//   Gr = sG(NL, 0.25) sG(NV, 0.25)
// Sampling: a lobe with probability proportional to (1 - metallic) (diffuse:
// cosine-weighted), 1 (specular: GGX half vector) and clearcoat / 4 (GTR1 half
// vector); the pdf of a direction is the mixture of the three.
template <class R> struct Disney {
  V3<R> base;
  R metallic, subsurface, specular, roughness, sheen, sheenTint, clearcoat, clearcoatGloss;
};
template <class R> static inline R sqr(R x) { return x * x; }
template <class R> static inline R schlick_w(R c) { R m = (R)1 - c; m = m < (R)0 ? (R)0 : (m > (R)1 ? (R)1 : m); return (m * m) * (m * m) * m; }
template <class R> static inline R gtr1(R nh, R a) {
  if (a >= (R)1) return (R)(1.0 / kPi);
  R a2 = a * a, t = (R)1 + (a2 - (R)1) * nh * nh;
  return (a2 - (R)1) / ((R)kPi * std::log(a2) * t);
}
template <class R> static inline R gtr2(R nh, R a) {
  R a2 = a * a, t = (R)1 + (a2 - (R)1) * nh * nh;
  return a2 / ((R)kPi * t * t);
}
template <class R> static inline R smith_g(R n, R a) { R a2 = a * a, b = n * n; return (R)1 / (n + std::sqrt(a2 + b - a2 * b)); }
template <class R> static inline R disney_alpha(const Disney<R>& m) { R a = m.roughness * m.roughness; return a < (R)1e-3 ? (R)1e-3 : a; }
template <class R> static inline R disney_cc_alpha(const Disney<R>& m) { return lerp((R)0.1, (R)0.001, m.clearcoatGloss); }
template <class R> static inline void disney_lobes(const Disney<R>& m, R& pd, R& ps, R& pc) {
  R wd = (R)1 - m.metallic, ws = (R)1, wc = (R)0.25 * m.clearcoat;
  R t = (wd + ws) + wc;
  pd = wd / t; ps = ws / t; pc = wc / t;
}
template <class R>
static V3<R> disney_f(const Disney<R>& m, V3<R> n, V3<R> wo, V3<R> wi) {
  R nl = dot(n, wi), nv = dot(n, wo);
  if (nl <= (R)0 || nv <= (R)0) return mk((R)0, (R)0, (R)0);
  V3<R> hv = wi + wo;
  R hl = length(hv);
  if (!(hl > (R)0)) return mk((R)0, (R)0, (R)0);
  V3<R> h = hv / hl;
  R nh = dot(n, h), lh = dot(wi, h);
  R lum = (R)0.3 * m.base.x + (R)0.6 * m.base.y + (R)0.1 * m.base.z;
  V3<R> tint = lum > (R)0 ? m.base / lum : mk((R)1, (R)1, (R)1);
  V3<R> cspec0 = vlerp(mk((R)1, (R)1, (R)1) * ((R)0.08 * m.specular), m.base, m.metallic);
  V3<R> csheen = vlerp(mk((R)1, (R)1, (R)1), tint, m.sheenTint);
  R fl = schlick_w(nl), fv = schlick_w(nv), fh = schlick_w(lh);
  R fd90 = (R)0.5 + (R)2 * lh * lh * m.roughness;
  R fd = ((R)1 + (fd90 - (R)1) * fl) * ((R)1 + (fd90 - (R)1) * fv);
  R fss90 = lh * lh * m.roughness;
  R fss = ((R)1 + (fss90 - (R)1) * fl) * ((R)1 + (fss90 - (R)1) * fv);
  R ss = (R)1.25 * (fss * ((R)1 / (nl + nv) - (R)0.5) + (R)0.5);
  R a = disney_alpha(m);
  R ds = gtr2(nh, a);
  V3<R> fs = vlerp(cspec0, mk((R)1, (R)1, (R)1), fh);
  R gs = smith_g(nl, a) * smith_g(nv, a);
  V3<R> fsheen = csheen * (fh * m.sheen);
  R dr = gtr1(nh, disney_cc_alpha(m));
  R fr = lerp((R)0.04, (R)1, fh);
  R gr = smith_g(nl, (R)0.25) * smith_g(nv, (R)0.25);
  V3<R> diffuse = (m.base * ((R)(1.0 / kPi) * lerp(fd, ss, m.subsurface)) + fsheen) * ((R)1 - m.metallic);
  return (diffuse + fs * (gs * ds)) + mk((R)1, (R)1, (R)1) * ((R)0.25 * m.clearcoat * gr * fr * dr);
}
template <class R>
static R disney_pdf(const Disney<R>& m, V3<R> n, V3<R> wo, V3<R> wi) {
  R nl = dot(n, wi), nv = dot(n, wo);
  if (nl <= (R)0 || nv <= (R)0) return (R)0;
  V3<R> hv = wi + wo;
  R hl = length(hv);
  if (!(hl > (R)0)) return (R)0;
  V3<R> h = hv / hl;
  R nh = dot(n, h), vh = dot(wo, h);
  if (vh <= (R)0) return (R)0;
  R pd, ps, pc;
  disney_lobes(m, pd, ps, pc);
  R pdf = pd * nl * (R)(1.0 / kPi);
  pdf = pdf + ps * gtr2(nh, disney_alpha(m)) * nh / ((R)4 * vh);
  pdf = pdf + pc * gtr1(nh, disney_cc_alpha(m)) * nh / ((R)4 * vh);
  return pdf;
}
// Direction by lobe choice u0 and (u1, u2); false when it leaves below the surface.
template <class R>
static bool disney_sample(const Disney<R>& m, V3<R> n, V3<R> wo, R u0, R u1, R u2, V3<R>& wi) {
  R pd, ps, pc;
  disney_lobes(m, pd, ps, pc);
  V3<R> u, v; basis(n, u, v);
  R phi = (R)(2.0 * kPi * (double)u1);
  if (u0 < pd) {                                   // cosine-weighted hemisphere
    R ct = std::sqrt(u2), st = std::sqrt((R)1 - u2);
    R sp, cp;
    sincos_r(phi, sp, cp);
    wi = normalize((n * ct + (u * st) * cp) + (v * st) * sp);
  } else {
    R a = u0 < pd + ps ? disney_alpha(m) : disney_cc_alpha(m);
    R ct;
    if (u0 < pd + ps) ct = std::sqrt(((R)1 - u2) / ((R)1 + (a * a - (R)1) * u2));                       // GGX
    else ct = a >= (R)1 ? std::sqrt((R)1 - u2)
                        : std::sqrt(((R)1 - std::pow(a * a, (R)1 - u2)) / ((R)1 - a * a));                // GTR1
    R st = std::sqrt(std::fmax((R)0, (R)1 - ct * ct));
    R sp, cp;
    sincos_r(phi, sp, cp);
    V3<R> h = (n * ct + (u * st) * cp) + (v * st) * sp;
    wi = normalize(h * ((R)2 * dot(wo, h)) - wo);
  }
  return dot(n, wi) > (R)0;
}

// ---------------------------------------------------------------- BDPT
struct RenderParams {
  u32 width, height;
  float cam_pos[3], cam_dir[3], cam_up[3], fov;
  u32 eye_depth, light_depth, light_paths;
  u32 flags;      // bit 0: literal shadow endpoint dist - eps (ill-conditioned, see shadow_max)
  u64 seed;
  u32 row0, rows; // rows [row0, row0 + rows) of the frame are traced (rows = 0: all)
  // adaptive mode (cudaRTBDPTStreamAdapBilatAtomic.cu:26-38)
  u32 mode;        // 0 plain (cudaRTBDPTStream), 1 adaptive proportional (AdapBilatAtomic),
                   // 2 rounds with bisection (AdapBilatBinary), 3 empirical variance rounds (AdapEmp),
                   // 4 path tracing (cudaRTPTStream), 5 with regeneration (cudaRTPTStreamRegen),
                   // 6 as 5 through the reference's own loop (render_pt_literal),
                   // 7 the Disney-BRDF path tracer with next-event estimation and MIS (extension),
                   // 8 the megakernel BDPT with camera-path restarts (cudaRTBDPTRegen, "Path Regen")
  u32 nd;          // DesiredSamplingN
  float sc, sp, sn, sd;   // filter colour / position / normal / diffuse sigmas
  u32 radius;      // filter radius
  u32 filter_diff; // Filter Diff Flag
  u32 warmup;      // plain frames before adaptive sampling starts (5)
  u32 pass_frames; // adaptive frames allocated together: pass_frames * W * H samples from one
                   // error map, traced under the pass's first frame (0 or 1: one frame a pass)
  float min_prob;  // MinTraceProb (rounds)
  u32 pad3;
  // extensions with no reference counterpart, 0 = off (the reference's behaviour)
  float ss;        // spatial sigma of the filter table (0: the colour sigma, AdapBilatAtomic.cu:1327)
  float sz;        // relative-depth sigma: (z_q - z_p) / z_p, z = |first-hit position - camera|
  u32 min_paths;   // warm-up frames raised to this many
  u32 max_paths;   // a pixel with this many eye paths gets no more
};

// Shadow-segment end for a connection.  The reference passes dist - eps
// (cudaRTBDPTStream.cu:685-686) but IntersectTri also measures t from the
// origin shifted by eps, so the target vertex's own surface lands on maxDist to
// within rounding and visibility becomes a coin flip of the arithmetic.  The
// defined endpoint stops one eps before the target in unshifted distance.
static inline double shadow_max(double dist, u32 flags) { return (flags & 1u) ? dist - kBiasEps : dist - 2.0 * kBiasEps; }

template <class R> struct LightVertex {
  V3<R> pos, norm, irrad, irradDir, diff, emissive;
  R specular, metallic, roughness;
};

// Shading inputs at a hit (cudaRTBDPTStream.cu:374-401).
template <class R> struct HitShade {
  V3<R> norm, pos, diff, emissive;
  R trans, specular, metallic, roughness;
};

// tex2D<float4> of the reference's texture objects (src/cudaRTCommon.cu:416-441):
// normalized coordinates wrapped to [0, 1), linear filtering between the four
// texels around (x - 1/2, y - 1/2) with wrapped indices, blended in x then y.
// NaN and infinite coordinates sample as 0.  CUDA's texture unit computes the
// blend weights in 9-bit fixed point; this is the exact-weight form.
template <class R>
static void tex_sample(const Scene& s, u32 id, R u, R v, R out[4]) {
  const u32 first = s.texDesc[id * 3], Wu = s.texDesc[id * 3 + 1], Hu = s.texDesc[id * 3 + 2];
  const int W = (int)Wu, H = (int)Hu;
  if (!(std::fabs(u) < (R)1e30)) u = 0;
  if (!(std::fabs(v) < (R)1e30)) v = 0;
  const R fu = u - std::floor(u), fv = v - std::floor(v);
  const R x = fu * (R)W - (R)0.5, y = fv * (R)H - (R)0.5;
  const R fx = std::floor(x), fy = std::floor(y);
  const R a = x - fx, b = y - fy;
  int i0 = std::min(std::max((int)fx, -1), W - 1), j0 = std::min(std::max((int)fy, -1), H - 1);
  if (i0 < 0) i0 += W;
  if (j0 < 0) j0 += H;
  const int i1 = i0 + 1 == W ? 0 : i0 + 1, j1 = j0 + 1 == H ? 0 : j0 + 1;
  const float* t00 = &s.tex[(first + (size_t)j0 * W + i0) * 4];
  const float* t10 = &s.tex[(first + (size_t)j0 * W + i1) * 4];
  const float* t01 = &s.tex[(first + (size_t)j1 * W + i0) * 4];
  const float* t11 = &s.tex[(first + (size_t)j1 * W + i1) * 4];
  const R ia = (R)1 - a, ib = (R)1 - b;
  for (int c = 0; c < 4; c++)
    out[c] = ((R)t00[c] * ia + (R)t10[c] * a) * ib + ((R)t01[c] * ia + (R)t11[c] * a) * b;
}

// Environment light (no reference counterpart): latitude-longitude map, u =
// atan2(d.z, d.x) / 2pi + 1/2, v = acos(d.y) / pi, bilinear between the four
// texel centres around (u W - 1/2, v H - 1/2), u wrapping, v clamped to the
// first and last rows; a non-finite direction sees 0.
template <class R>
static V3<R> env_radiance(const Scene& s, V3<R> d) {
  const int W = (int)s.envW, H = (int)s.envH;
  const R u = atan2_r(d.z, d.x) * (R)(0.5 / kPi) + (R)0.5;
  const R v = acos_r(std::fmin(std::fmax(d.y, (R)-1), (R)1)) * (R)(1.0 / kPi);
  const R x = u * (R)W - (R)0.5, y = v * (R)H - (R)0.5;
  if (!(std::fabs(x) < (R)1e30) || !(std::fabs(y) < (R)1e30)) return mk((R)0, (R)0, (R)0);
  const R fx = std::floor(x), fy = std::floor(y);
  const R a = x - fx, b = y - fy;
  int i0 = (int)fx % W;
  if (i0 < 0) i0 += W;
  const int i1 = i0 + 1 == W ? 0 : i0 + 1;
  const int j0 = std::min(std::max((int)fy, 0), H - 1), j1 = std::min(std::max((int)fy + 1, 0), H - 1);
  const float* t00 = &s.env[((size_t)j0 * W + i0) * 3];
  const float* t10 = &s.env[((size_t)j0 * W + i1) * 3];
  const float* t01 = &s.env[((size_t)j1 * W + i0) * 3];
  const float* t11 = &s.env[((size_t)j1 * W + i1) * 3];
  const R ia = (R)1 - a, ib = (R)1 - b;
  R c[3];
  for (int k = 0; k < 3; k++)
    c[k] = ((R)t00[k] * ia + (R)t10[k] * a) * ib + ((R)t01[k] * ia + (R)t11[k] * a) * b;
  return mk(c[0], c[1], c[2]);
}

template <class R>
static HitShade<R> fetch_hit(const Scene& s, u32 leaf, R w, R u, R v) {
  HitShade<R> hs;
  const u32* t = &s.tri[leaf * 3];
  const Material& m = s.mats[s.tmat[leaf]];
  V3<R> n0 = cvt<R>(&s.nrm[t[0] * 3]), n1 = cvt<R>(&s.nrm[t[1] * 3]), n2 = cvt<R>(&s.nrm[t[2] * 3]);
  V3<R> q0 = cvt<R>(&s.pos[t[0] * 3]), q1 = cvt<R>(&s.pos[t[1] * 3]), q2 = cvt<R>(&s.pos[t[2] * 3]);
  hs.norm = (n0 * w + n1 * u) + n2 * v;
  hs.pos = (q0 * w + q1 * u) + q2 * v;
  hs.diff = cvt<R>(m.diffuse); hs.emissive = cvt<R>(m.emissive);
  hs.trans = m.transparency; hs.specular = m.specular; hs.metallic = m.metallic; hs.roughness = m.roughness;
  // GetMaterialColors (src/cudaRTCommon.cu:369-406); uv = uv0 w + uv1 u + uv2 v
  // (cudaRTBDPTStream.cu:379-382); ids outside the textures mean none
  const u32 ntex = (u32)(s.texDesc.size() / 3);
  const bool dtex = (u32)m.diffuseTexId < ntex, etex = (u32)m.emissiveTexId < ntex;
  if (dtex || etex) {
    const R tu = ((R)s.uv[t[0] * 2] * w + (R)s.uv[t[1] * 2] * u) + (R)s.uv[t[2] * 2] * v;
    const R tv = ((R)s.uv[t[0] * 2 + 1] * w + (R)s.uv[t[1] * 2 + 1] * u) + (R)s.uv[t[2] * 2 + 1] * v;
    R c[4];
    if (dtex) {
      tex_sample<R>(s, (u32)m.diffuseTexId, tu, tv, c);
      hs.diff = mk(c[0], c[1], c[2]);
      hs.trans = c[3] * (R)m.transparency;
    }
    if (etex) {
      tex_sample<R>(s, (u32)m.emissiveTexId, tu, tv, c);
      hs.emissive = mk(c[0], c[1], c[2]);
    }
  }
  return hs;
}

enum { RT_EYE = 0, RT_DIFF = 1, RT_SPEC = 2, RT_LIGHT = 3 };

// One BRDF sampling step shared by eye and light bounces
// (cudaRTBDPTStream.cu:228-329 and :402-502).
template <class R> struct Bounce { V3<R> orig, dir, mul; int type; };
template <class R>
// tir_reflect: the path tracers turn a refraction past the critical angle into
// a reflection (cudaRTPTStream.cu:152-160); the BDPT files do not.
static Bounce<R> sample_bounce(Pcg32& rng, V3<R> ro, V3<R> rd, R t, const HitShade<R>& hs, V3<R> nl, int curType,
                               bool refl_offset = false, bool tir_reflect = false) {
  R x0 = uniform<R>(rng), x1 = uniform<R>(rng);
  V3<R> hDir = sample_ggx(x0, x1, hs.roughness, nl);
  V3<R> ndir = rd * (R)-1;
  R voH = dot(ndir, hDir);
  V3<R> f0 = vlerp(mk((R)0.08 * hs.specular, (R)0.08 * hs.specular, (R)0.08 * hs.specular), hs.diff, hs.metallic);
  V3<R> F = f_schlick(f0, voH);
  R reflProb = lerp(length(F), (R)1, hs.metallic);
  R refrProb = hs.trans;
  V3<R> reflDir = mk((R)0, (R)0, (R)0), refrDir = mk((R)0, (R)0, (R)0);
  Bounce<R> b; b.orig = ro; b.dir = rd; b.type = curType; b.mul = mk((R)0, (R)0, (R)0);
  if (refrProb > (R)0) {
    bool into = dot(nl, hs.norm) > (R)0;
    R nt = hs.specular * (R)0.8 + (R)1, nc = (R)1;
    R nnt = into ? nc / nt : nt / nc;
    R ddn = dot(hDir, rd);
    R cos2t = (R)1 - nnt * nnt * ((R)1 - ddn * ddn);
    if (tir_reflect && cos2t < (R)0) reflProb = (R)1;
    else refrDir = normalize(rd * nnt - hDir * (ddn * nnt + std::sqrt(cos2t)));
  }
  if (reflProb > (R)0) {
    reflDir = normalize(rd - (hDir * (R)2) * dot(hDir, rd));
    if (dot(reflDir, nl) < (R)0) reflProb = (R)0;
  }
  if (prob_rand(rng, reflProb)) {
    b.orig = ro + rd * (R)((double)t - kBiasEps); b.dir = reflDir;
    // the adaptive files also step off along the reflection (AdapBilatAtomic.cu:584)
    if (refl_offset) b.orig = b.orig + reflDir * (R)kBiasEps;
    R VoH = dot(ndir, hDir), NoV = dot(nl, ndir), NoH = dot(nl, hDir), NoL = dot(nl, reflDir);
    R G = vis_smith(hs.roughness, NoV, NoL);
    b.mul = (F * G * VoH) / (NoV * NoH * reflProb);
    b.type = RT_SPEC;
  } else if (prob_rand(rng, refrProb)) {
    b.orig = (ro + rd * (R)((double)t + kBiasEps)) + refrDir * (R)kBiasEps; b.dir = refrDir;
    R cosine = dot(nl * (R)-1, refrDir);
    b.mul = hs.diff * cosine / (refrProb * ((R)1 - reflProb));
    b.type = RT_SPEC;
  } else {
    V3<R> u, v; basis(nl, u, v);
    R r1 = (R)(2.0 * kPi * (double)uniform<R>(rng));
    R r2cos = std::sqrt(uniform<R>(rng));
    R r2sin = (R)1 - r2cos * r2cos;
    R s1, c1;
    sincos_r(r1, s1, c1);
    V3<R> dd = normalize((nl * r2cos + (u * r2sin) * c1) + (v * r2sin) * s1);
    b.orig = (ro + rd * t) + dd * (R)kBiasEps; b.dir = dd;
    b.mul = lambert(hs.diff) * (R)kPi / (((R)1 - refrProb) * ((R)1 - reflProb));
    b.type = RT_DIFF;
  }
  return b;
}

template <class R> struct Camera { V3<R> pos, dir, up, right; R fov; };

template <class R>
static Camera<R> make_camera(const RenderParams& p) {
  // cudaRTBDPTStream.cu:823-824 (host float Vec3 math)
  F3 d = f3(p.cam_dir[0], p.cam_dir[1], p.cam_dir[2]), up = f3(p.cam_up[0], p.cam_up[1], p.cam_up[2]);
  F3 r = f3(d.y * up.z - d.z * up.y, d.z * up.x - d.x * up.z, d.x * up.y - d.y * up.x);
  float rl = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z); r = f3(r.x / rl, r.y / rl, r.z / rl);
  F3 u2 = f3(r.y * d.z - r.z * d.y, r.z * d.x - r.x * d.z, r.x * d.y - r.y * d.x);
  float ul = std::sqrt(u2.x * u2.x + u2.y * u2.y + u2.z * u2.z); u2 = f3(u2.x / ul, u2.y / ul, u2.z / ul);
  Camera<R> c;
  c.pos = mk((R)p.cam_pos[0], (R)p.cam_pos[1], (R)p.cam_pos[2]);
  c.dir = mk((R)d.x, (R)d.y, (R)d.z); c.up = mk((R)u2.x, (R)u2.y, (R)u2.z); c.right = mk((R)r.x, (R)r.y, (R)r.z);
  c.fov = (R)p.fov;
  return c;
}

// cudaRTBDPTStream.cu:629-637
template <class R>
static V3<R> eye_dir(const Camera<R>& c, u32 x, u32 y, R W, R H, Pcg32& rng) {
  R tf = std::tan(c.fov * (R)0.5);
  R u = ((((R)2 * ((R)x + (R)0.5)) / W - (R)1) * tf * W) / H;
  R v = (((R)2 * ((R)y + (R)0.5)) / H - (R)1) * tf;
  R au = u + ((uniform<R>(rng) - (R)0.5) / H) * tf;
  R av = v + ((uniform<R>(rng) - (R)0.5) / H) * tf;
  return normalize((c.right * au + c.up * av) + c.dir);
}

template <class R> struct FrameState {
  std::vector<R> result;      // W*H*3 running mean
  std::vector<u32> N;         // W*H sample counts
  std::vector<R> gpos, gnorm, gdiff;   // W*H*3 first-hit running means (adaptive mode)
  std::vector<R> sq;                   // W*H*3 running mean of squared samples (squareResult)
  u32 rounds = 0;                      // last round-based allocation: rounds run
  std::vector<float> filtered, err;    // last adaptive frame: filter output, error map
  std::vector<u32> counts;             // last adaptive frame: samples per pixel
  std::vector<u32> paths;              // eye paths traced per pixel since frame 0
  float err_sum = 0.f;
  u64 ext_rays = 0, shadow_rays = 0;
  std::vector<LightVertex<R>> lv;
  u32 lvN = 0;
  u32 light_group = 0xFFFFFFFFu;       // three-frame group whose cache lv holds
};

// Light-vertex cache pass (cudaRTBDPTStream.cu:180-363,545-617,875-906).
template <class R>
static void light_pass(const Scene& s, const RenderParams& p, u32 frame, FrameState<R>& fs) {
  u32 L = p.light_paths;
  fs.lv.assign((size_t)L * (p.light_depth + 1), LightVertex<R>());
  fs.lvN = 0;
  if (s.lightTri.empty()) return;
  struct LP { Pcg32 rng; V3<R> orig, dir, sample, mul; int type; u32 depth, origin; bool term; };
  std::vector<LP> lp(L);
  u32 nlt = (u32)s.lightTri.size();
  for (u32 x = 0; x < L; x++) {
    LP& P = lp[x];
    pcg32_seed(P.rng, stream_key(p.seed, frame, DOMAIN_LIGHT, 0), x);
    u32 id = (u32)(uniform<R>(P.rng) * (R)nlt); if (id >= nlt) id = nlt - 1;
    R lw = uniform<R>(P.rng), lu = uniform<R>(P.rng);
    if (lw + lu > (R)1) { lw = (R)1 - lw; lu = (R)1 - lu; }
    R lvv = ((R)1 - lw) - lu;
    HitShade<R> hs = fetch_hit<R>(s, s.lightTri[id], lw, lu, lvv);
    V3<R> u, v; basis(hs.norm, u, v);
    R r1 = (R)(2.0 * kPi * (double)uniform<R>(P.rng));
    R r2cos = std::sqrt(uniform<R>(P.rng));
    R r2sin = (R)1 - r2cos * r2cos;
    R s1, c1;
    sincos_r(r1, s1, c1);
    V3<R> dd = normalize((hs.norm * r2cos + (u * r2sin) * c1) + (v * r2sin) * s1);
    P.orig = hs.pos + hs.norm * (R)kBiasEps; P.dir = dd; P.sample = hs.emissive; P.mul = mk((R)1, (R)1, (R)1);
    P.type = RT_LIGHT; P.depth = 0; P.origin = x; P.term = false;
    LightVertex<R>& V = fs.lv[x];
    V.irrad = hs.emissive; V.irradDir = mk((R)0, (R)0, (R)0); V.norm = hs.norm; V.pos = hs.pos;
    V.diff = hs.diff; V.emissive = hs.emissive; V.specular = hs.specular; V.metallic = hs.metallic; V.roughness = hs.roughness;
  }
  u32 cur = L;
  for (;;) {
    // stream position of each live path = rank among live paths (stable compaction)
    u32 rank = 0;
    for (u32 x = 0; x < L; x++) {
      LP& P = lp[x];
      if (P.term) continue;
      u32 slot = cur + rank++;
      HitR<R> h;
      bool hit = trace_defined<R>(s, P.orig, P.dir, (R)kInf, (R)kBiasEps, false, h);
      if (!hit) { fs.lv[slot] = fs.lv[P.origin]; P.term = true; continue; }
      R w = ((R)1 - h.u) - h.v;
      HitShade<R> hs = fetch_hit<R>(s, h.leaf, w, h.u, h.v);
      V3<R> nl = dot(hs.norm, P.dir) < (R)0 ? hs.norm : hs.norm * (R)-1;
      LightVertex<R>& V = fs.lv[slot];
      V.irrad = P.sample; V.irradDir = P.dir * (R)-1; V.norm = nl; V.pos = hs.pos;
      V.diff = hs.diff; V.emissive = hs.emissive; V.specular = hs.specular; V.metallic = hs.metallic; V.roughness = hs.roughness;
      Bounce<R> b = sample_bounce<R>(P.rng, P.orig, P.dir, h.t, hs, nl, P.type);
      if (b.type != RT_DIFF) V.irrad = mk((R)0, (R)0, (R)0);
      if (dot(b.dir, nl) < (R)0) V.norm = V.norm * (R)-1;
      P.sample = hs.emissive + vmul(P.sample, b.mul);
      R contrib = length(P.mul) * length(b.mul);
      if (length(hs.emissive) > (R)0) contrib = (R)0;
      if (uniform<R>(P.rng) > contrib || P.depth + 1 >= p.light_depth) P.term = true;
      else { P.mul = vmul(P.mul, b.mul); P.dir = b.dir; P.orig = b.orig; P.type = b.type; P.depth++; }
    }
    if (rank == 0) break;
    cur += rank;
  }
  fs.lvN = cur;
}

template <class R> struct PathResult { V3<R> accum; u32 N; V3<R> gpos, gnorm, gdiff; u32 ext_rays, shadow_rays; };

// One eye path (cudaRTBDPTStream.cu:365-543,665-723).
template <class R>
static PathResult<R> eye_path(const Scene& s, const RenderParams& p, const Camera<R>& cam, const FrameState<R>& fs,
                              u32 frame, u32 x, u32 y, u32 sample, int variant) {
  u32 W = p.width, H = p.height, pix = y * W + x;
  Pcg32 rng; pcg32_seed(rng, stream_key(p.seed, frame, DOMAIN_EYE, sample), pix);
  V3<R> dir = eye_dir(cam, x, y, (R)W, (R)H, rng);
  V3<R> orig = cam.pos, mul = mk((R)1, (R)1, (R)1), psample = mk((R)0, (R)0, (R)0);
  PathResult<R> pr; pr.accum = mk((R)0, (R)0, (R)0); pr.N = 0;
  pr.gpos = pr.gnorm = pr.gdiff = mk((R)0, (R)0, (R)0);
  pr.ext_rays = pr.shadow_rays = 0;
  for (u32 depth = 0;; ) {
    HitR<R> h;
    pr.ext_rays++;
    if (!trace_defined<R>(s, orig, dir, (R)kInf, (R)kBiasEps, false, h)) {
      // an escaping eye path gathers the environment's radiance through its throughput
      if (s.envW) psample = psample + vmul(env_radiance<R>(s, dir), mul);
      pr.accum = pr.accum + psample; pr.N++; break;
    }
    R w = ((R)1 - h.u) - h.v;
    HitShade<R> hs = fetch_hit<R>(s, h.leaf, w, h.u, h.v);
    V3<R> nl = dot(hs.norm, dir) < (R)0 ? hs.norm : hs.norm * (R)-1;
    if (depth == 0) { pr.gpos = hs.pos; pr.gnorm = nl; pr.gdiff = hs.diff; }
    Bounce<R> b = sample_bounce<R>(rng, orig, dir, h.t, hs, nl, RT_EYE, variant == 1);
    psample = psample + vmul(hs.emissive, mul);
    V3<R> inMul = mul, inDir = dir * (R)-1;
    R contrib = length(mul) * length(b.mul);
    if (length(hs.emissive) > (R)0) contrib = (R)0;
    if (uniform<R>(rng) > contrib || depth + 1 >= p.eye_depth) {
      pr.accum = pr.accum + psample; pr.N++; break;
    }
    mul = vmul(mul, b.mul); dir = b.dir; depth++;
    orig = b.orig;
    if (b.type == RT_DIFF && fs.lvN > 0) {
      // pt_connectEyeLightPath_kernel + GetLightFromRandLightVertices
      V3<R> pos = orig + nl * (R)kBiasEps;
      u32 li = (u32)(uniform<R>(rng) * (R)fs.lvN); if (li >= fs.lvN) li = fs.lvN - 1;
      const LightVertex<R>& L = fs.lv[li];
      V3<R> irrad = mk((R)0, (R)0, (R)0), irradDir = mk((R)0, (R)0, (R)0);
      V3<R> toL = normalize(L.pos - pos);
      R dist = length(L.pos - pos);
      if (length(L.irrad) > (R)0 && dot(nl, toL) > (R)0) {
        HitR<R> sh;
        pr.shadow_rays++;
        if (!trace_defined<R>(s, pos, toL, (R)shadow_max((double)dist, p.flags), (R)kBiasEps, false, sh, true)) {
          if ((double)dist > kFltEps) {
            irrad = shading(toL * (R)-1, L.irradDir, L.irrad, L.norm, L.diff, L.metallic, L.roughness, L.specular, (R)1, (R)1) + L.emissive;
            irradDir = toL;
          } else { irrad = L.irrad; irradDir = L.irradDir * (R)-1; }
        }
      }
      V3<R> c = vmax(mk((R)0, (R)0, (R)0),
                     shading(inDir, irradDir, irrad, nl, hs.diff, hs.metallic, hs.roughness, hs.specular, (R)1 - hs.trans, (R)1));
      if (length(c) > (R)0) { pr.accum = pr.accum + vmul(c, inMul); pr.N += 4; }
    }
  }
  return pr;
}

// pt_applyPathQueueResult_kernel (cudaRTBDPTStream.cu:741-778)
template <class R>
static inline void apply_sample(R* res, u32& Nold, V3<R> accum, u32 n) {
  if (n == 0) return;
  u32 nn = Nold + n;
  if (!(nn > Nold) || isinf3(accum)) return;
  R ri = (R)1 / (R)nn, oi = (R)Nold * ri;
  res[0] = std::fmax(ri * accum.x + oi * res[0], (R)0);
  res[1] = std::fmax(ri * accum.y + oi * res[1], (R)0);
  res[2] = std::fmax(ri * accum.z + oi * res[2], (R)0);
  Nold = nn;
}

template <class R>
static void render_frames(const Scene& s, const RenderParams& p, u32 frame0, u32 nframes, FrameState<R>& fs,
                          R* gbuf /* optional W*H*9: pos, norm, diff of the last frame */) {
  Camera<R> cam = make_camera<R>(p);
  u32 W = p.width, H = p.height;
  for (u32 f = frame0; f < frame0 + nframes; f++) {
    if (f % 3 == 0) light_pass<R>(s, p, f, fs);
    if (f == 0) std::fill(fs.N.begin(), fs.N.end(), 0u);
    u64 er = 0, sr = 0;
    const u32 y_end = std::min<u32>(H, p.rows ? p.row0 + p.rows : H);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : er, sr)
    for (i32 y = (i32)std::min(p.row0, H); y < (i32)y_end; y++) {
      for (u32 x = 0; x < W; x++) {
        PathResult<R> pr = eye_path<R>(s, p, cam, fs, f, x, (u32)y, 0, 0);
        er += pr.ext_rays; sr += pr.shadow_rays;
        u32 pix = (u32)y * W + x;
        apply_sample(&fs.result[pix * 3], fs.N[pix], pr.accum, pr.N);
        if (gbuf) {
          R* g = gbuf + (size_t)pix * 9;
          g[0] = pr.gpos.x; g[1] = pr.gpos.y; g[2] = pr.gpos.z;
          g[3] = pr.gnorm.x; g[4] = pr.gnorm.y; g[5] = pr.gnorm.z;
          g[6] = pr.gdiff.x; g[7] = pr.gdiff.y; g[8] = pr.gdiff.z;
        }
      }
    }
    fs.ext_rays += er; fs.shadow_rays += sr;
  }
}

// ---------------------------------------------------------------- path tracers
// One path of cudaRTPTStream / cudaRTPTStreamRegen (PTPathVertex,
// cudaRTPTStream.cu:22-48): pt_begin draws the eye ray from the path's RNG
// (pt_genPathQueue_kernel :266-284, pt_reinitRegenPathStream_kernel
// cudaRTPTStreamRegen.cu:306-326, which keep accum and N), pt_step is one
// bounce of pt_traceSample_kernel (cudaRTPTStream.cu:91-264): emission
// gathered, BRDF sample with total internal reflection turned into a
// reflection, Russian roulette, no connection.  regen: the diffuse-to-specular
// cut of cudaRTPTStreamRegen.cu:241.
template <class R> struct PtPath {
  Pcg32 rng;
  V3<R> orig, dir, mul, psample, accum;
  u32 N = 0, depth = 0, x = 0, y = 0;
  int type = RT_EYE;
  bool ended = false;
};
template <class R>
static void pt_begin(const RenderParams& p, const Camera<R>& cam, PtPath<R>& P) {
  P.dir = eye_dir(cam, P.x, P.y, (R)p.width, (R)p.height, P.rng);
  P.orig = cam.pos;
  P.mul = mk((R)1, (R)1, (R)1);
  P.psample = mk((R)0, (R)0, (R)0);
  P.depth = 0;
  P.type = RT_EYE;
  P.ended = false;
}
// Returns true when the path ends (its sample added to accum, N).
template <class R>
static bool pt_step(const Scene& s, PtPath<R>& P, u32 limit, bool regen, u64& rays) {
  HitR<R> h;
  rays++;
  if (!trace_defined<R>(s, P.orig, P.dir, (R)kInf, (R)kBiasEps, false, h)) {
    if (s.envW) P.psample = P.psample + vmul(env_radiance<R>(s, P.dir), P.mul);
    P.accum = P.accum + P.psample; P.N++; P.ended = true;
    return true;
  }
  R w = ((R)1 - h.u) - h.v;
  HitShade<R> hs = fetch_hit<R>(s, h.leaf, w, h.u, h.v);
  V3<R> nl = dot(hs.norm, P.dir) < (R)0 ? hs.norm : hs.norm * (R)-1;
  Bounce<R> b = sample_bounce<R>(P.rng, P.orig, P.dir, h.t, hs, nl, RT_EYE, false, true);
  P.psample = P.psample + vmul(hs.emissive, P.mul);
  R contrib = length(P.mul) * length(b.mul);
  if (length(hs.emissive) > (R)0) contrib = (R)0;
  if (regen && P.type == RT_DIFF && b.type == RT_SPEC) contrib = (R)0;
  if (uniform<R>(P.rng) > contrib || P.depth + 1 >= limit) {
    P.accum = P.accum + P.psample; P.N++; P.ended = true;
    return true;
  }
  P.mul = vmul(P.mul, b.mul); P.dir = b.dir; P.depth++;
  P.orig = b.orig; P.type = b.type;
  return false;
}

// ---------------------------------------------------------------- Disney path tracer (extension)
// Mode 7, no reference counterpart: a unidirectional path tracer with the
// Disney BRDF above and next-event estimation, each emissive-triangle hit
// weighted against the light sample by the power heuristic (Veach 1997,
// section 9.2.4).  Light sampling: an emissive triangle uniformly among the
// scene's (the reference's leaf order), a uniform point on it; emitters emit
// on both faces, as the reference's do.  Per bounce the stream gives, in
// order: the Russian-roulette uniform (bounce >= 3; survival
// min(0.95, max T), survivors reweighted), the light sample (three uniforms,
// when the scene has emissive triangles) and the BRDF sample (three).  A path
// has at most eye_depth segments, the light sample's included.  Surfaces are
// opaque (transparency is not read).  The environment light, when set, is
// gathered by escaping paths only, with weight 1 (it is never light-sampled).
// flags bits of the mode's estimator checks: light samples off (emitter hits
// weighted 1), emitter hits off past the first segment (light samples
// weighted 1), Russian roulette off; every variant has the same expectation.
constexpr u32 kDisneyBsdfOnly = 2u, kDisneyLightOnly = 4u, kDisneyNoRoulette = 8u;
template <class R>
static Disney<R> disney_of(const Scene& s, u32 leaf, const HitShade<R>& hs) {
  const Material& mt = s.mats[s.tmat[leaf]];
  Disney<R> m;
  m.base = hs.diff; m.metallic = hs.metallic; m.specular = hs.specular; m.roughness = hs.roughness;
  m.subsurface = mt.subsurface; m.sheen = mt.sheen; m.sheenTint = mt.sheenTint;
  m.clearcoat = mt.clearcoat; m.clearcoatGloss = mt.clearcoatGloss;
  return m;
}
// Solid-angle density of sampling the point at distance t along unit d on
// leaf triangle `leaf` by the light sampler.
template <class R>
static R light_pdf(const Scene& s, u32 leaf, R t, V3<R> d, u32 nL) {
  V3<R> ng = cross(cvt<R>(&s.e0[leaf * 3]), cvt<R>(&s.e1[leaf * 3]));
  R a2 = length(ng);
  R c = std::fabs(dot(ng, d)) / a2;
  if (!(c > (R)0) || !(a2 > (R)0)) return (R)0;
  return t * t / (c * (a2 * (R)0.5) * (R)nL);
}
template <class R>
static void disney_path(const Scene& s, const RenderParams& p, const Camera<R>& cam, u32 frame, u32 x, u32 y,
                        V3<R>& L, u64& rays, u64& srays) {
  const u32 W = p.width, H = p.height, pix = y * W + x;
  Pcg32 rng; pcg32_seed(rng, stream_key(p.seed, frame, DOMAIN_EYE, 0), pix);
  V3<R> dir = eye_dir(cam, x, y, (R)W, (R)H, rng), orig = cam.pos, T = mk((R)1, (R)1, (R)1);
  L = mk((R)0, (R)0, (R)0);
  R prev_pdf = (R)0;
  const u32 nL = (u32)s.lightTri.size();
  for (u32 d = 0;; d++) {
    HitR<R> h;
    rays++;
    if (!trace_defined<R>(s, orig, dir, (R)kInf, (R)kBiasEps, false, h)) {
      if (s.envW) L = L + vmul(T, env_radiance<R>(s, dir));
      return;
    }
    R w = ((R)1 - h.u) - h.v;
    HitShade<R> hs = fetch_hit<R>(s, h.leaf, w, h.u, h.v);
    if (length(hs.emissive) > (R)0) {
      R wt = (R)1;
      if (d > 0 && !(p.flags & kDisneyBsdfOnly)) {
        R pl = light_pdf<R>(s, h.leaf, h.t, dir, nL);
        wt = (p.flags & kDisneyLightOnly) ? (R)0 : prev_pdf * prev_pdf / (prev_pdf * prev_pdf + pl * pl);
      }
      L = L + vmul(T, hs.emissive) * wt;
    }
    if (d + 1 >= p.eye_depth) return;
    if (d >= 3 && !(p.flags & kDisneyNoRoulette)) {
      R q = std::fmin((R)0.95, std::fmax(T.x, std::fmax(T.y, T.z)));
      if (uniform<R>(rng) >= q) return;
      T = T / q;
    }
    V3<R> nl = dot(hs.norm, dir) < (R)0 ? hs.norm : hs.norm * (R)-1;
    V3<R> pos = orig + dir * h.t, wo = dir * (R)-1;
    Disney<R> m = disney_of<R>(s, h.leaf, hs);
    if (nL > 0 && !(p.flags & kDisneyBsdfOnly)) {
      R l0 = uniform<R>(rng), l1 = uniform<R>(rng), l2 = uniform<R>(rng);
      u32 li = (u32)(l0 * (R)nL); if (li >= nL) li = nL - 1;
      if (l1 + l2 > (R)1) { l1 = (R)1 - l1; l2 = (R)1 - l2; }
      const u32 lt = s.lightTri[li];
      V3<R> e0 = cvt<R>(&s.e0[lt * 3]), e1 = cvt<R>(&s.e1[lt * 3]);
      V3<R> lp = (cvt<R>(&s.p0[lt * 3]) + e0 * l1) + e1 * l2;
      V3<R> so = pos + nl * (R)kBiasEps;
      V3<R> toL = lp - so;
      R dist = length(toL);
      if ((double)dist > kFltEps) {
        V3<R> wi = toL / dist;
        R cs = dot(nl, wi);
        R pl = light_pdf<R>(s, lt, dist, wi, nL);
        if (cs > (R)0 && pl > (R)0) {
          V3<R> f = disney_f(m, nl, wo, wi);
          R pb = disney_pdf(m, nl, wo, wi);
          R wt = (p.flags & kDisneyLightOnly) ? (R)1 : pl * pl / (pl * pl + pb * pb);
          const Material& lm = s.mats[s.tmat[lt]];
          V3<R> c = vmul(vmul(T, f), cvt<R>(lm.emissive)) * (cs * wt / pl);
          if (length(c) > (R)0) {
            HitR<R> sh;
            srays++;
            if (!trace_defined<R>(s, so, wi, (R)((double)dist - 2.0 * kBiasEps), (R)kBiasEps, false, sh, true))
              L = L + c;
          }
        }
      }
    }
    R b0 = uniform<R>(rng), b1 = uniform<R>(rng), b2 = uniform<R>(rng);
    V3<R> wi;
    if (!disney_sample(m, nl, wo, b0, b1, b2, wi)) return;
    R pdf = disney_pdf(m, nl, wo, wi);
    if (!(pdf > (R)0)) return;
    T = vmul(T, disney_f(m, nl, wo, wi)) * (dot(nl, wi) / pdf);
    prev_pdf = pdf;
    orig = pos + wi * (R)kBiasEps;
    dir = wi;
  }
}
template <class R>
static void render_disney(const Scene& s, const RenderParams& p, u32 frame0, u32 nframes, FrameState<R>& fs) {
  Camera<R> cam = make_camera<R>(p);
  const u32 W = p.width, H = p.height;
  for (u32 f = frame0; f < frame0 + nframes; f++) {
    if (f == 0) std::fill(fs.N.begin(), fs.N.end(), 0u);
    u64 er = 0, sr = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : er, sr)
    for (i32 y = 0; y < (i32)H; y++)
      for (u32 x = 0; x < W; x++) {
        V3<R> Lr;
        disney_path<R>(s, p, cam, f, x, (u32)y, Lr, er, sr);
        const u32 pix = (u32)y * W + x;
        apply_sample(&fs.result[pix * 3], fs.N[pix], Lr, 1u);
      }
    fs.ext_rays += er; fs.shadow_rays += sr;
  }
}

// ---------------------------------------------------------------- "Path Regen" BDPT (cudaRTBDPTRegen)
// The reference's megakernel BDPT renderer (src/cudaRTBDPTRegen.cu), one
// thread per pixel, restated without recursion:
//  - every third frame a light pass fills a 640-vertex cache: 128 light paths
//    of 5 vertices each (bdpt_light_kernel :314-330, bdpt_light_ray :80-309).
//    A path starts on a random emissive triangle (its vertex: irrad =
//    emissive, no direction) and then carries the throughput (1, 1, 1), not the
//    emission (:179, :307); each later vertex stores the throughput it arrived
//    with and the hit's interpolated normal; a miss stores an empty vertex and
//    the next vertex starts a new path, as does a throughput below 1e-4;
//  - the eye sample runs at most `eye_depth` levels of pt0_normalRay
//    (:433-625).  Per level: the BRDF bounce (refraction past the critical
//    angle reflects, :501-505), the pixel weight multiplied by |f|, set to 0
//    on an emitter; Russian roulette on it (:562): a failed test restarts a
//    camera path for the same pixel (count + 1), a passed one connects a
//    diffuse bounce to a uniformly chosen cache vertex (count + 4 when it
//    contributes).  A miss or the level bound adds 1 to the count.  Radiance
//    is summed back up the levels, r = f * r' + emission + connection, and
//    a level where a camera path starts banks r into the pixel's sum (:598-606).
static const u32 kRegenLightN = 640, kRegenLightDepth = 5, kRegenMaxLevels = 8;
enum { DOMAIN_REGEN_LIGHT = 4, DOMAIN_REGEN_EYE = 5 };

template <class R> static LightVertex<R> zero_vertex() {
  LightVertex<R> V;
  V.pos = V.norm = V.irrad = V.irradDir = V.diff = V.emissive = mk((R)0, (R)0, (R)0);
  V.specular = V.metallic = V.roughness = (R)0;
  return V;
}
template <class R> static LightVertex<R> shade_vertex(const HitShade<R>& hs, V3<R> irrad, V3<R> irradDir) {
  LightVertex<R> V;
  V.pos = hs.pos; V.norm = hs.norm; V.irrad = irrad; V.irradDir = irradDir;
  V.diff = hs.diff; V.emissive = hs.emissive; V.specular = hs.specular; V.metallic = hs.metallic;
  V.roughness = hs.roughness;
  return V;
}

template <class R>
static void regen_light_pass(const Scene& s, const RenderParams& p, u32 frame, FrameState<R>& fs) {
  fs.lv.assign(kRegenLightN, zero_vertex<R>());
  fs.lvN = kRegenLightN;
  const u32 nlt = (u32)s.lightTri.size();
  for (u32 x = 0; x < kRegenLightN / kRegenLightDepth; x++) {
    Pcg32 rng; pcg32_seed(rng, stream_key(p.seed, frame, DOMAIN_REGEN_LIGHT, 0), x);
    bool valid = false;
    V3<R> o = mk((R)0, (R)0, (R)0), d = o, irr = o;
    for (u32 k = 0; k < kRegenLightDepth; k++) {
      V3<R> mulT = mk((R)1, (R)1, (R)1), no = o, nd = d;
      bool nvalid = false;
      LightVertex<R> V = zero_vertex<R>();
      if (!valid || length(irr) < (R)kBiasEps) {
        if (nlt > 0) {                                   // a new path on an emitter (:94-158)
          u32 id = (u32)(uniform<R>(rng) * (R)nlt); if (id >= nlt) id = nlt - 1;
          R lw = uniform<R>(rng), lu = uniform<R>(rng);
          if (lw + lu > (R)1) { lw = (R)1 - lw; lu = (R)1 - lu; }
          R lvv = ((R)1 - lw) - lu;
          HitShade<R> hs = fetch_hit<R>(s, s.lightTri[id], lw, lu, lvv);
          V3<R> u, v; basis(hs.norm, u, v);
          R r1 = (R)(2.0 * kPi * (double)uniform<R>(rng));
          R r2cos = std::sqrt(uniform<R>(rng));
          R r2sin = (R)1 - r2cos * r2cos;
          R s1, c1;
          sincos_r(r1, s1, c1);
          nd = normalize((hs.norm * r2cos + (u * r2sin) * c1) + (v * r2sin) * s1);
          no = hs.pos + hs.norm * (R)kBiasEps;
          nvalid = true;
          V = shade_vertex(hs, hs.emissive, mk((R)0, (R)0, (R)0));
        }
      } else {
        HitR<R> h;
        if (trace_defined<R>(s, o, d, (R)kInf, (R)kBiasEps, false, h)) {   // :159-299
          R w = ((R)1 - h.u) - h.v;
          HitShade<R> hs = fetch_hit<R>(s, h.leaf, w, h.u, h.v);
          V3<R> nl = dot(hs.norm, d) < (R)0 ? hs.norm : hs.norm * (R)-1;
          Bounce<R> b = sample_bounce<R>(rng, o, d, h.t, hs, nl, RT_LIGHT, false, true);
          no = b.orig; nd = b.dir; nvalid = true;
          mulT = vmul(b.mul, irr);
          V = shade_vertex(hs, irr, d * (R)-1);
        }
      }
      fs.lv[x * kRegenLightDepth + k] = V;
      valid = nvalid; o = no; d = nd; irr = mulT;
    }
  }
}

template <class R> struct RegenSample { V3<R> accum; u32 N; u32 ext_rays, shadow_rays, paths; };

template <class R>
static RegenSample<R> regen_eye(const Scene& s, const RenderParams& p, const Camera<R>& cam, const FrameState<R>& fs,
                                u32 frame, u32 x, u32 y) {
  const u32 W = p.width, H = p.height, pix = y * W + x;
  const u32 levels = p.eye_depth < kRegenMaxLevels ? p.eye_depth : kRegenMaxLevels;
  Pcg32 rng; pcg32_seed(rng, stream_key(p.seed, frame, DOMAIN_REGEN_EYE, 0), pix);
  RegenSample<R> rs; rs.accum = mk((R)0, (R)0, (R)0); rs.N = 0; rs.ext_rays = rs.shadow_rays = 0; rs.paths = 1;
  const R tf = std::tan(cam.fov * (R)0.5), Wf = (R)W, Hf = (R)H;
  const R pu = ((((R)2 * ((R)x + (R)0.5)) / Wf - (R)1) * tf * Wf) / Hf;   // pt0_kernel :646-647
  const R pv = (((R)2 * ((R)y + (R)0.5)) / Hf - (R)1) * tf;
  auto camera_ray = [&]() {
    R au = pu + ((uniform<R>(rng) - (R)0.5) / Hf) * tf;
    R av = pv + ((uniform<R>(rng) - (R)0.5) / Hf) * tf;
    return normalize((cam.right * au + cam.up * av) + cam.dir);
  };
  V3<R> o = cam.pos, d = camera_ray();
  R weight = (R)1;
  u32 depth = 0, count = 0, D = levels;
  V3<R> mul[kRegenMaxLevels], em[kRegenMaxLevels], lc[kRegenMaxLevels];
  bool bank[kRegenMaxLevels];
  for (u32 L = 0; L < levels; L++) {
    HitR<R> h;
    rs.ext_rays++;
    if (!trace_defined<R>(s, o, d, (R)kInf, (R)kBiasEps, false, h)) { D = L; break; }   // :620-623
    R w = ((R)1 - h.u) - h.v;
    HitShade<R> hs = fetch_hit<R>(s, h.leaf, w, h.u, h.v);
    V3<R> nl = dot(hs.norm, d) < (R)0 ? hs.norm : hs.norm * (R)-1;
    Bounce<R> b = sample_bounce<R>(rng, o, d, h.t, hs, nl, RT_EYE, false, true);
    weight = weight * length(b.mul);
    if (length(hs.emissive) > (R)0) weight = (R)0;
    bank[L] = depth == 0;
    V3<R> c = mk((R)0, (R)0, (R)0), no = b.orig, nd = b.dir;
    if (!prob_rand(rng, weight)) {                       // a new camera path for the pixel (:562-571)
      weight = (R)1; depth = 0; count++;
      no = cam.pos; nd = camera_ray();
      rs.paths++;
    } else {
      depth++;
      if (b.type == RT_DIFF) {                            // GetLightFromRandLightVertices (:400-431)
        V3<R> P = hs.pos + nl * (R)kBiasEps;
        u32 li = (u32)(uniform<R>(rng) * (R)kRegenLightN); if (li >= kRegenLightN) li = kRegenLightN - 1;
        const LightVertex<R>& LV = fs.lv[li];
        V3<R> irrad = mk((R)0, (R)0, (R)0), irradDir = irrad;
        V3<R> toL = normalize(LV.pos - P);
        R dist = length(LV.pos - P);
        if (length(LV.irrad) > (R)0 && dot(nl, toL) > (R)0) {
          HitR<R> sh;
          rs.shadow_rays++;
          if (!trace_defined<R>(s, P, toL, (R)shadow_max((double)dist, p.flags), (R)kBiasEps, false, sh, true)) {
            if (length(LV.irradDir) > (R)kFltEps)
              irrad = shading(toL * (R)-1, LV.irradDir, LV.irrad, LV.norm, LV.diff, LV.metallic, LV.roughness,
                              LV.specular, (R)1, (R)1) + LV.emissive;
            else irrad = LV.irrad;
            irradDir = toL;
          }
        }
        c = vmax(mk((R)0, (R)0, (R)0), shading(d * (R)-1, irradDir, irrad, nl, hs.diff, hs.metallic, hs.roughness,
                                               hs.specular, (R)1 - hs.trans, (R)1));
        if (length(c) > (R)0) count += 4;
      }
    }
    mul[L] = b.mul; em[L] = hs.emissive; lc[L] = c;
    o = no; d = nd;
  }
  count++;                                                 // the miss or the level bound (:620-630)
  V3<R> r = mk((R)0, (R)0, (R)0);
  for (u32 k = D; k-- > 0;) {
    r = (vmul(mul[k], r) + em[k]) + lc[k];
    if (bank[k]) { rs.accum = rs.accum + r; r = mk((R)0, (R)0, (R)0); }
  }
  rs.N = count;
  return rs;
}

// pt0_kernel's running mean (:655-664): the count is every camera path started
// plus 4 per contributing connection, reset at frame 0.
template <class R>
static void render_bdpt_regen(const Scene& s, const RenderParams& p, u32 frame0, u32 nframes, FrameState<R>& fs) {
  Camera<R> cam = make_camera<R>(p);
  const u32 W = p.width, H = p.height;
  for (u32 f = frame0; f < frame0 + nframes; f++) {
    if (f % 3 == 0) regen_light_pass<R>(s, p, f, fs);
    if (f == 0) std::fill(fs.N.begin(), fs.N.end(), 0u);
    u64 er = 0, sr = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : er, sr)
    for (i32 y = 0; y < (i32)H; y++)
      for (u32 x = 0; x < W; x++) {
        const RegenSample<R> rs = regen_eye<R>(s, p, cam, fs, f, x, (u32)y);
        const u32 pix = (u32)y * W + x;
        R* res = &fs.result[pix * 3];
        const u32 nn = fs.N[pix] + rs.N;
        const R ri = (R)1 / (R)nn, oi = (R)fs.N[pix] * ri;
        res[0] = std::fmax(ri * rs.accum.x + oi * res[0], (R)0);
        res[1] = std::fmax(ri * rs.accum.y + oi * res[1], (R)0);
        res[2] = std::fmax(ri * rs.accum.z + oi * res[2], (R)0);
        fs.N[pix] = nn;
        er += rs.ext_rays; sr += rs.shadow_rays;
      }
    fs.ext_rays += er; fs.shadow_rays += sr;
  }
}

// pt_applyPathQueueResult_kernel (cudaRTPTStream.cu:303-325): the frame's
// per-pixel mean weighted 1 / (frame + 1).
template <class R>
static inline void apply_pt(R* res, u32& Nold, u32 frame, V3<R> accum, u32 n) {
  if (n > 0) {
    R fn = (R)n;
    R sx = accum.x / fn, sy = accum.y / fn, sz = accum.z / fn;
    R ri = (R)1 / (R)(frame + 1), oi = (R)1 - ri;
    res[0] = std::fmax(ri * sx + oi * res[0], (R)0);
    res[1] = std::fmax(ri * sy + oi * res[1], (R)0);
    res[2] = std::fmax(ri * sz + oi * res[2], (R)0);
  }
  Nold += n;
}

// Frames of the path tracers (mode 4: cudaRTPTStream::Render, :327-432; mode 5:
// cudaRTPTStreamRegen::Render, :374-478), all rows.  Regeneration: a pixel whose
// path ended at bounce b < T, T the frame's last bounce of a first path,
// restarts once with its RNG continuing and traces bounces b+1 .. T (depth
// bound min(eye depth, T - b)); a restart still running at T adds what it
// gathered, as an ended one does.  That is the reference's schedule while its
// stream holds the whole frame (W*H <= 640,000); render_pt_literal runs that
// schedule itself.
template <class R>
static void render_pt(const Scene& s, const RenderParams& p, u32 frame0, u32 nframes, FrameState<R>& fs) {
  Camera<R> cam = make_camera<R>(p);
  const u32 W = p.width, H = p.height, n = W * H;
  const bool regen = p.mode == 5;
  std::vector<PtPath<R>> P(n);
  std::vector<u32> term(n);
  for (u32 f = frame0; f < frame0 + nframes; f++) {
    if (f == 0) std::fill(fs.N.begin(), fs.N.end(), 0u);
    u64 er = 0;
    u32 T = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : er) reduction(max : T)
    for (i32 y = 0; y < (i32)H; y++) {
      for (u32 x = 0; x < W; x++) {
        const u32 pix = (u32)y * W + x;
        PtPath<R>& A = P[pix];
        A = PtPath<R>();
        A.x = x; A.y = (u32)y;
        pcg32_seed(A.rng, stream_key(p.seed, f, DOMAIN_EYE, 0), pix);
        A.accum = mk((R)0, (R)0, (R)0);
        pt_begin(p, cam, A);
        while (!pt_step(s, A, p.eye_depth, regen, er)) {
        }
        term[pix] = A.depth;
        T = std::max(T, A.depth);
      }
    }
    if (regen) {
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : er)
      for (i32 y = 0; y < (i32)H; y++) {
        for (u32 x = 0; x < W; x++) {
          const u32 pix = (u32)y * W + x;
          if (term[pix] >= T) continue;
          PtPath<R>& A = P[pix];
          pt_begin(p, cam, A);
          while (!pt_step(s, A, std::min(T - term[pix], p.eye_depth), true, er)) {
          }
        }
      }
    }
    for (u32 pix = 0; pix < n; pix++) apply_pt(&fs.result[pix * 3], fs.N[pix], f, P[pix].accum, P[pix].N);
    fs.ext_rays += er;
  }
}

// cudaRTPTStreamRegen::Render's loop as written (:424-471), for a frame the
// stream holds whole: per iteration the restarts queued by the previous one
// begin, every live first path and restart takes one bounce, ended restarts
// leave the list for good, ended first paths queue their restart; the loop
// stops when no first path is live, and a restart still live adds its partial
// sample in pt_applyPathQueueResult_kernel.  Single-threaded.
template <class R>
static void render_pt_literal(const Scene& s, const RenderParams& p, u32 frame0, u32 nframes, FrameState<R>& fs) {
  Camera<R> cam = make_camera<R>(p);
  const u32 W = p.width, H = p.height, n = W * H;
  std::vector<PtPath<R>> P(n);
  for (u32 f = frame0; f < frame0 + nframes; f++) {
    if (f == 0) std::fill(fs.N.begin(), fs.N.end(), 0u);
    u64 er = 0;
    std::vector<u32> stream, regen_live, regen_queued;
    for (u32 pix = 0; pix < n; pix++) {
      PtPath<R>& A = P[pix];
      A = PtPath<R>();
      A.x = pix % W; A.y = pix / W;
      pcg32_seed(A.rng, stream_key(p.seed, f, DOMAIN_EYE, 0), pix);
      A.accum = mk((R)0, (R)0, (R)0);
      pt_begin(p, cam, A);
      stream.push_back(pix);
    }
    while (!stream.empty()) {
      for (u32 pix : regen_queued) { pt_begin(p, cam, P[pix]); regen_live.push_back(pix); }
      regen_queued.clear();
      for (u32 pix : stream) pt_step(s, P[pix], p.eye_depth, true, er);
      for (u32 pix : regen_live) pt_step(s, P[pix], p.eye_depth, true, er);
      std::vector<u32> keep;
      for (u32 pix : regen_live) if (!P[pix].ended) keep.push_back(pix);
      regen_live.swap(keep);
      keep.clear();
      for (u32 pix : stream) (P[pix].ended ? regen_queued : keep).push_back(pix);
      stream.swap(keep);
    }
    for (u32 pix : regen_live) { P[pix].accum = P[pix].accum + P[pix].psample; P[pix].N++; }
    for (u32 pix = 0; pix < n; pix++) apply_pt(&fs.result[pix * 3], fs.N[pix], f, P[pix].accum, P[pix].N);
    fs.ext_rays += er;
  }
}

// ---------------------------------------------------------------- filter
// cudaBilateralFilter.cu:5-24
static void gaussian_table(float sigma, u32 radius, std::vector<float>& g) {
  g.assign(2 * radius + 1, 0.f);
  for (u32 i = 0; i < 2 * radius + 1; i++) {
    int x = (int)i - (int)radius;
    g[i] = std::exp(-(float)(x * x) / (2 * sigma * sigma));
  }
}
// The reference clamps with CUDA's mixed int/unsigned min/max: (x + j) is
// unsigned, so a negative coordinate wraps and clamps to the HIGH edge.
static inline u32 ref_clamp(u32 x, int j, u32 n) {
  u32 v = x + (u32)j;          // unsigned wrap as in the reference
  return v < n - 1 ? v : n - 1;
}
template <class R>
static inline R euclid(const float* a, const float* b, R d) {
  R m = ((R)(b[0] - a[0]) * (R)(b[0] - a[0]) + (R)(b[1] - a[1]) * (R)(b[1] - a[1])) + (R)(b[2] - a[2]) * (R)(b[2] - a[2]);
  return std::exp(-m / ((R)2 * d * d));
}
// bilaterial_posnorm_kernel / _posnormemit_kernel (cudaBilateralFilter.cu:66-142)
// Depth of a first-hit position: its distance from the camera.
template <class R> static inline R depth_of(const float* p, const float* cam) {
  R dx = (R)(p[0] - cam[0]), dy = (R)(p[1] - cam[1]), dz = (R)(p[2] - cam[2]);
  return std::sqrt((dx * dx + dy * dy) + dz * dz);
}
template <class R>
static void bilateral(const float* color, const float* pos, const float* norm, const float* emit, u32 W, u32 H,
                      float sc, float sp, float sn, float se, u32 radius, const std::vector<float>& g, float* out,
                      float sz = 0.f, const float* cam = nullptr) {
#pragma omp parallel for schedule(static)
  for (i32 yy = 0; yy < (i32)H; yy++) {
    u32 y = (u32)yy;
    for (u32 x = 0; x < W; x++) {
      u32 ind = x + y * W;
      R sum = 0, t0 = 0, t1 = 0, t2 = 0;
      const R zc = sz > 0.f ? depth_of<R>(&pos[ind * 3], cam) : (R)0;
      for (int i = -(int)radius; i <= (int)radius; i++) {
        for (int j = -(int)radius; j <= (int)radius; j++) {
          u32 c = ref_clamp(x, j, W) + ref_clamp(y, i, H) * W;
          R f = (R)g[i + radius] * (R)g[j + radius];
          f = f * euclid<R>(&color[c * 3], &color[ind * 3], (R)sc);
          f = f * euclid<R>(&pos[c * 3], &pos[ind * 3], (R)sp);
          f = f * euclid<R>(&norm[c * 3], &norm[ind * 3], (R)sn);
          if (emit) f = f * euclid<R>(&emit[c * 3], &emit[ind * 3], (R)se);
          if (sz > 0.f) {
            const R dz = zc > (R)0 ? (depth_of<R>(&pos[c * 3], cam) - zc) / zc : (R)0;
            f = f * std::exp(-(dz * dz) / ((R)2 * (R)sz * (R)sz));
          }
          t0 += f * (R)color[c * 3]; t1 += f * (R)color[c * 3 + 1]; t2 += f * (R)color[c * 3 + 2];
          sum += f;
        }
      }
      out[ind * 3] = (float)(t0 / sum); out[ind * 3 + 1] = (float)(t1 / sum); out[ind * 3 + 2] = (float)(t2 / sum);
    }
  }
}

// ---------------------------------------------------------------- defined filter arithmetic
// e^x by Cody-Waite argument reduction (Cody & Waite, "Software Manual for the
// Elementary Functions", 1980) and a degree-7 Taylor polynomial on
// |r| <= ln2/2, scaled by 2^k with ldexp.  The device kernel evaluates the same
// fma sequence, so filter weights agree bit for bit between the two.
static inline float cpt_expf(float x) {
  if (x < -103.0f) return 0.f;
  if (x > 88.0f) return std::numeric_limits<float>::infinity();
  float k = std::nearbyint(x * 1.44269504f);
  float r = std::fma(k, -0.693145751953125f, x);
  r = std::fma(k, -1.42860676533018e-06f, r);
  float p = 1.f / 5040.f;
  p = std::fma(p, r, 1.f / 720.f);
  p = std::fma(p, r, 1.f / 120.f);
  p = std::fma(p, r, 1.f / 24.f);
  p = std::fma(p, r, 1.f / 6.f);
  p = std::fma(p, r, 0.5f);
  p = std::fma(p, r, 1.f);
  p = std::fma(p, r, 1.f);
  return std::ldexp(p, (int)k);
}
static inline float sqdist(const float* a, const float* b) {
  float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return (dx * dx + dy * dy) + dz * dz;
}
// |a - b|^2 of a range term, as the filter kernel evaluates it:
// fma(dz, dz, fma(dy, dy, dx * dx)).
static inline float sqd_fma(const float* a, const float* b) {
  const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}
// Cross-bilateral with one exponential per tap:
// w = (G[i] G[j]) * exp(-(|dc|^2/2sc^2 + |dp|^2/2sp^2 + |dn|^2/2sn^2 [+ |dd|^2/2sd^2])),
// the product of the reference's per-term exponentials (cudaBilateralFilter.cu:89-92)
// folded into one; FP64 runs keep the literal product (bilateral<double>).
// The optional depth term adds ((z_q - z_p) / z_p)^2 / 2sz^2 last, z the
// correctly rounded distance of the first hit from the camera.
// The arithmetic is the device kernel's: each 1 / 2s^2 and 1 / z_p correctly
// rounded once, the exponent accumulated with fmas in the order above
// (c * (1/2sc^2), then fma per further term), the colour sums with fmas.
static void bilateral_defined(const float* color, const float* pos, const float* norm, const float* emit, u32 W, u32 H,
                              float sc, float sp, float sn, float se, u32 radius, const std::vector<float>& g,
                              float* out, float sz = 0.f, const float* cam = nullptr) {
  const float kc = 2.f * sc * sc, kp = 2.f * sp * sp, kn = 2.f * sn * sn, ke = 2.f * se * se, kz = 2.f * sz * sz;
  const float ikc = 1.f / kc, ikp = 1.f / kp, ikn = 1.f / kn, ike = 1.f / ke, ikz = kz > 0.f ? 1.f / kz : 0.f;
#pragma omp parallel for schedule(static)
  for (i32 yy = 0; yy < (i32)H; yy++) {
    u32 y = (u32)yy;
    for (u32 x = 0; x < W; x++) {
      u32 ind = x + y * W;
      float sum = 0.f, t0 = 0.f, t1 = 0.f, t2 = 0.f;
      const float zc = kz > 0.f ? std::sqrt(sqdist(&pos[ind * 3], cam)) : 0.f;
      const float izc = zc > 0.f ? 1.f / zc : 0.f;
      for (int i = -(int)radius; i <= (int)radius; i++) {
        for (int j = -(int)radius; j <= (int)radius; j++) {
          u32 c = ref_clamp(x, j, W) + ref_clamp(y, i, H) * W;
          float a = sqd_fma(&color[c * 3], &color[ind * 3]) * ikc;
          a = std::fma(sqd_fma(&pos[c * 3], &pos[ind * 3]), ikp, a);
          a = std::fma(sqd_fma(&norm[c * 3], &norm[ind * 3]), ikn, a);
          if (emit) a = std::fma(sqd_fma(&emit[c * 3], &emit[ind * 3]), ike, a);
          if (kz > 0.f) {
            const float zq = std::sqrt(sqdist(&pos[c * 3], cam));
            const float dz = (zq - zc) * izc;
            a = std::fma(dz * dz, ikz, a);
          }
          float f = (g[i + radius] * g[j + radius]) * cpt_expf(-a);
          t0 = std::fma(f, color[c * 3], t0);
          t1 = std::fma(f, color[c * 3 + 1], t1);
          t2 = std::fma(f, color[c * 3 + 2], t2);
          sum = sum + f;
        }
      }
      out[ind * 3] = t0 / sum; out[ind * 3 + 1] = t1 / sum; out[ind * 3 + 2] = t2 / sum;
    }
  }
}

// Fixed-order double sum: pairwise tree inside 256-element blocks, blocks in
// order.  Stands in for thrust::reduce (AdapBilatAtomic.cu:1444), whose float
// order is unspecified; the device reduction uses the same tree.
// The device's fixed-order sum (k_err_block_sums, k_err_final): a pairwise
// tree over each 256-element block, the same tree over each chunk of 256 block
// sums, then the chunk sums in order.
static double tree256(const double* v, size_t n) {
  double s[256];
  for (size_t t = 0; t < 256; t++) s[t] = t < n ? v[t] : 0.0;
  for (int st = 128; st >= 1; st >>= 1)
    for (int t = 0; t < st; t++) s[t] += s[t + st];
  return s[0];
}
static double tree_sum(const float* e, size_t n) {
  std::vector<double> bs;
  double blk[256];
  for (size_t b = 0; b < n; b += 256) {
    const size_t m = std::min<size_t>(256, n - b);
    for (size_t t = 0; t < m; t++) blk[t] = (double)e[b + t];
    bs.push_back(tree256(blk, m));
  }
  double total = 0.0;
  for (size_t c = 0; c < bs.size(); c += 256) total += tree256(bs.data() + c, std::min<size_t>(256, bs.size() - c));
  return total;
}

// pt_genAdapPathQueue_kernel count (AdapBilatAtomic.cu:830-831)
static inline u32 prop_count(float e, float sum, u32 gen) {
  float r = (float)gen * (e / sum);
  if (!(r >= 0.f) || !(r < 4294967040.f)) return 0;
  return (u32)r;
}

// pt_applyPathQueueResult_kernel of the adaptive file (AdapBilatAtomic.cu:984-1039):
// running means of radiance and of the count-weighted first-hit G-buffer.
template <class R>
static inline void apply_sample_gbuf(FrameState<R>& fs, u32 pix, const PathResult<R>& pr) {
  u32& Nold = fs.N[pix];
  u32 nn = Nold + pr.N;
  if (pr.N == 0 || !(nn > Nold) || isinf3(pr.accum)) return;
  R ri = (R)1 / (R)nn, oi = (R)Nold * ri;
  R* res = &fs.result[pix * 3];
  res[0] = std::fmax(ri * pr.accum.x + oi * res[0], (R)0);
  res[1] = std::fmax(ri * pr.accum.y + oi * res[1], (R)0);
  res[2] = std::fmax(ri * pr.accum.z + oi * res[2], (R)0);
  Nold = nn;
  R n = (R)pr.N;
  auto upd = [&](std::vector<R>& buf, V3<R> v) {
    V3<R> w = v * n;
    R* q = &buf[pix * 3];
    q[0] = ri * w.x + oi * q[0]; q[1] = ri * w.y + oi * q[1]; q[2] = ri * w.z + oi * q[2];
  };
  upd(fs.gpos, pr.gpos); upd(fs.gnorm, pr.gnorm); upd(fs.gdiff, pr.gdiff);
  // squareResult of pt_applyPathQueueResult_kernel (AdapEmp.cu:927-944)
  R* q = &fs.sq[pix * 3];
  q[0] = std::fmax(ri * (pr.accum.x * pr.accum.x) + oi * q[0], (R)0);
  q[1] = std::fmax(ri * (pr.accum.y * pr.accum.y) + oi * q[1], (R)0);
  q[2] = std::fmax(ri * (pr.accum.z * pr.accum.z) + oi * q[2], (R)0);
}

// Round-based allocation: pt_genTempAdapPathQueue_kernel rounds, remove_if,
// pt_fillTempAdapPathQueue_kernel (AdapBilatBinary.cu:1348-1388 with the window
// bisection; AdapEmp.cu:1232-1260 with the window at max e).  Round r keeps
// pixel x unless u * window > max(e_x, minProb), u from the PCG32 stream of
// (frame, slots filled so far) at pixel x; the first kept pixels in pixel order
// take the remaining slots; unfilled slots go to pixels 0, 1, ... (wrapping).
static const u32 kMaxRounds = 1024;
static void alloc_rounds(const RenderParams& p, u32 frame, const std::vector<float>& err, u32 n, u32 gen,
                         bool bisect, float desired, float min_prob, std::vector<u32>& counts, u32& rounds) {
  float maxe = 0.f;                                   // thrust::maximum reduce from 0
  for (u32 x = 0; x < n; x++) maxe = std::max(maxe, std::fmax(err[x], 0.f));
  float left = 0.f, right = maxe * 2.f;
  float window = bisect ? (left + right) * 0.5f : maxe;
  counts.assign(n, 0u);
  std::vector<unsigned char> keep(n);
  u32 accum = 0;
  rounds = 0;
  while (accum < gen && rounds < kMaxRounds) {
    u32 kept = 0;
    for (u32 x = 0; x < n; x++) {
      Pcg32 rng; pcg32_seed(rng, stream_key(p.seed, frame, 3, accum), x);
      float u = uniform<float>(rng);
      keep[x] = (u * window > std::fmax(err[x], min_prob)) ? 0 : 1;
      kept += keep[x];
    }
    u32 take = std::min(kept, gen - accum), rank = 0;
    for (u32 x = 0; x < n; x++)
      if (keep[x]) { if (rank < take) counts[x]++; rank++; }
    accum += take;
    rounds++;
    if (take == 0) break;
    if (bisect) {
      if ((float)take < desired) right = window;
      else if ((float)take > desired) left = window;
      window = (left + right) * 0.5f;
    }
  }
  u32 unfilled = gen > accum ? gen - accum : 0u;
  for (u32 x = 0; x < n; x++) counts[x] += unfilled / n + (x < unfilled % n ? 1u : 0u);
}

// One adaptive frame after warm-up (AdapBilatAtomic.cu:1407-1538).
template <class R>
static void adaptive_frame(const Scene& s, const RenderParams& p, const Camera<R>& cam, u32 f, u32 gen,
                           FrameState<R>& fs) {
  const u32 W = p.width, H = p.height, n = W * H;
  std::vector<float> col(n * 3), pos(n * 3), nrm(n * 3), dif(n * 3);
  for (size_t i = 0; i < (size_t)n * 3; i++) {
    col[i] = (float)fs.result[i]; pos[i] = (float)fs.gpos[i]; nrm[i] = (float)fs.gnorm[i]; dif[i] = (float)fs.gdiff[i];
  }
  fs.err.assign(n, 0.f);
  fs.counts.assign(n, 0u);
  if (p.mode == 3) {
    // pt_calculateEmpiricalSquareError_kernel (AdapEmp.cu:949-960)
    fs.filtered.clear();
    for (u32 x = 0; x < n; x++) {
      u32 sn = fs.N[x];
      const R* a = &fs.result[x * 3];
      const R* b = &fs.sq[x * 3];
      R e = ((b[0] - a[0] * a[0]) + (b[1] - a[1] * a[1]) + (b[2] - a[2] * a[2])) * (R)sn / ((R)(u32)(sn - 1u) * (R)3);
      fs.err[x] = std::fmin((float)e, 1.0f);
    }
  } else {
    std::vector<float> g;
    gaussian_table(p.ss > 0.f ? p.ss : p.sc, p.radius, g);   // spatial sigma = colour sigma (:1327)
    fs.filtered.assign(n * 3, 0.f);
    const float* emit = p.filter_diff ? dif.data() : nullptr;
    if (sizeof(R) == 4) {
      bilateral_defined(col.data(), pos.data(), nrm.data(), emit, W, H, p.sc, p.sp, p.sn, p.sd, p.radius, g,
                        fs.filtered.data(), p.sz, p.cam_pos);
    } else {
      bilateral<double>(col.data(), pos.data(), nrm.data(), emit, W, H, p.sc, p.sp, p.sn, p.sd, p.radius, g,
                        fs.filtered.data(), p.sz, p.cam_pos);
    }
    // pt_calculateSquareError_kernel (:1143-1156; AdapBilatBinary.cu:1031-1044)
    for (u32 x = 0; x < n; x++) {
      u32 sn = fs.N[x];
      float red = (float)p.nd / (float)(p.nd + sn);
      if (p.nd + sn < sn) red = 0;
      const float* a = &fs.filtered[x * 3];
      const float* b = &col[x * 3];
      float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
      fs.err[x] = red * ((d0 * d0 + d1 * d1) + d2 * d2) / 3.f;
    }
  }
  // per-pixel cap: a pixel that has max_paths eye paths gets no weight, and its
  // count stops at max_paths - paths
  auto room = [&](u32 x) { return fs.paths[x] < p.max_paths ? p.max_paths - fs.paths[x] : 0u; };
  if (p.max_paths)
    for (u32 x = 0; x < n; x++)
      if (fs.paths[x] >= p.max_paths) fs.err[x] = 0.f;
  if (p.mode == 1) {
    float sum = (float)tree_sum(fs.err.data(), n);
    fs.err_sum = sum;
    for (u32 x = 0; x < n; x++) fs.counts[x] = prop_count(fs.err[x], sum, gen);
  } else {
    alloc_rounds(p, f, fs.err, n, gen, p.mode == 2, (float)n / (float)(p.nd ? p.nd : 1u), p.min_prob, fs.counts,
                 fs.rounds);
  }
  if (p.max_paths)
    for (u32 x = 0; x < n; x++) fs.counts[x] = std::min(fs.counts[x], room(x));
  for (u32 x = 0; x < n; x++) fs.paths[x] += fs.counts[x];
  const int variant = p.mode == 1 ? 1 : 0;      // reflection offset of AdapBilatAtomic's eye bounce
  u64 er = 0, sr = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : er, sr)
  for (i32 ii = 0; ii < (i32)n; ii++) {
    u32 pix = (u32)ii, y = pix / W, x = pix - y * W;
    V3<R> ts = mk((R)0, (R)0, (R)0), tp = ts, tn = ts, td = ts, tq = ts;
    u32 tN = 0;
    for (u32 k = 0; k < fs.counts[pix]; k++) {
      PathResult<R> pr = eye_path<R>(s, p, cam, fs, f, x, y, k, variant);
      er += pr.ext_rays; sr += pr.shadow_rays;
      if (pr.N == 0 || isinf3(pr.accum)) continue;        // pt_atomicApplyPathQueueResult_kernel (:1056-1061)
      R nf = (R)pr.N;
      ts = ts + pr.accum; tN += pr.N;
      tq = tq + vmul(pr.accum, pr.accum);
      tp = tp + pr.gpos * nf; tn = tn + pr.gnorm * nf; td = td + pr.gdiff * nf;
    }
    // pt_accumTempResultToResult_kernel (:1117-1140)
    u32 Nold = fs.N[pix], nn = tN + Nold;
    if (nn > Nold) {
      R ri = (R)1 / (R)nn, oi = (R)Nold * ri;
      R* res = &fs.result[pix * 3];
      res[0] = std::fmax(ri * ts.x + oi * res[0], (R)0);
      res[1] = std::fmax(ri * ts.y + oi * res[1], (R)0);
      res[2] = std::fmax(ri * ts.z + oi * res[2], (R)0);
      fs.N[pix] = Nold + tN;
      auto upd = [&](std::vector<R>& buf, V3<R> v) {
        R* q = &buf[pix * 3];
        q[0] = ri * v.x + oi * q[0]; q[1] = ri * v.y + oi * q[1]; q[2] = ri * v.z + oi * q[2];
      };
      upd(fs.gpos, tp); upd(fs.gnorm, tn); upd(fs.gdiff, td);
      R* q = &fs.sq[pix * 3];
      q[0] = std::fmax(ri * tq.x + oi * q[0], (R)0);
      q[1] = std::fmax(ri * tq.y + oi * q[1], (R)0);
      q[2] = std::fmax(ri * tq.z + oi * q[2], (R)0);
    }
  }
  fs.ext_rays += er; fs.shadow_rays += sr;
}

template <class R>
static void render_adaptive(const Scene& s, const RenderParams& p, u32 frame0, u32 nframes, FrameState<R>& fs) {
  Camera<R> cam = make_camera<R>(p);
  const u32 W = p.width, H = p.height, n = W * H;
  if (fs.gpos.size() != (size_t)n * 3) {
    fs.gpos.assign(n * 3, (R)0); fs.gnorm.assign(n * 3, (R)0); fs.gdiff.assign(n * 3, (R)0);
    fs.sq.assign(n * 3, (R)0);
    fs.paths.assign(n, 0u);
  }
  const int variant = p.mode == 1 ? 1 : 0;
  const u32 warmup = std::max(p.warmup, p.min_paths);       // every pixel gets min_paths paths
  for (u32 f = frame0; f < frame0 + nframes;) {
    // the cache of the frame's three-frame group (built at its first frame)
    if (f % 3 == 0 || fs.light_group != f / 3) light_pass<R>(s, p, f - f % 3, fs);
    fs.light_group = f / 3;
    u32 k = 1;
    if (f < warmup) {
      if (f == 0) { std::fill(fs.N.begin(), fs.N.end(), 0u); std::fill(fs.paths.begin(), fs.paths.end(), 0u); }
      for (u32 x = 0; x < n; x++) fs.paths[x]++;
      u64 er = 0, sr = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : er, sr)
      for (i32 ii = 0; ii < (i32)n; ii++) {
        u32 pix = (u32)ii, y = pix / W, x = pix - y * W;
        PathResult<R> pr = eye_path<R>(s, p, cam, fs, f, x, y, 0, variant);
        er += pr.ext_rays; sr += pr.shadow_rays;
        apply_sample_gbuf(fs, pix, pr);
      }
      fs.ext_rays += er; fs.shadow_rays += sr;
    } else {
      k = std::min(std::max(p.pass_frames, 1u), frame0 + nframes - f);
      adaptive_frame<R>(s, p, cam, f, k * n, fs);
    }
    f += k;
  }
}

}  // namespace orc

// ====================================================================== C ABI
using namespace orc;

template <class R> static V3<R> v3d(const double* p) { return mk((R)p[0], (R)p[1], (R)p[2]); }
template <class R> static void put3(V3<R> a, double* o) { o[0] = (double)a.x; o[1] = (double)a.y; o[2] = (double)a.z; }
template <class R>
static void bounce_many(const double* ro, const double* rd, double t, const double* hs, const double* nl, u64 seed,
                        u32 n, int refl_offset, i32* type, double* orig, double* dir, double* mul) {
  HitShade<R> h;
  h.norm = v3d<R>(hs); h.pos = mk((R)0, (R)0, (R)0); h.diff = v3d<R>(hs + 3); h.emissive = mk((R)0, (R)0, (R)0);
  h.trans = (R)hs[6]; h.specular = (R)hs[7]; h.metallic = (R)hs[8]; h.roughness = (R)hs[9];
  for (u32 i = 0; i < n; i++) {
    // independent streams: PCG32 streams that share an initial state and differ only in
    // the increment give related first draws, so each sample gets its own hashed state
    Pcg32 rng; pcg32_seed(rng, splitmix64_mix(seed + (u64)i * 0x9E3779B97F4A7C15ull), i);
    Bounce<R> b = sample_bounce<R>(rng, v3d<R>(ro), v3d<R>(rd), (R)t, h, v3d<R>(nl), RT_EYE, refl_offset != 0);
    type[i] = b.type;
    put3(b.orig, orig + (size_t)i * 3); put3(b.dir, dir + (size_t)i * 3); put3(b.mul, mul + (size_t)i * 3);
  }
}

extern "C" {

u32 orc_pcg32_stream(u64 initstate, u64 initseq, u32 n, u32* out) {
  Pcg32 r; pcg32_seed(r, initstate, initseq);
  for (u32 i = 0; i < n; i++) out[i] = pcg32_next(r);
  return n;
}
u64 orc_splitmix64_next(u64* state) {
  *state += 0x9E3779B97F4A7C15ULL;
  return splitmix64_mix(*state);
}
u64 orc_stream_key(u64 seed, u32 frame, u32 domain, u32 sample) { return stream_key(seed, frame, domain, sample); }
float orc_det_eps() { return det_eps<float>(); }

void* orc_scene_create(const float* pos3, const float* nrm3, u32 nv, const u32* idx3, const u32* mat, u32 nt,
                       const void* materials, u32 nmat, u32 max_depth) {
  Scene* s = new Scene();
  s->nv = nv; s->nt = nt;
  s->uv.assign((size_t)nv * 2, 0.f);
  s->pos.assign(pos3, pos3 + (size_t)nv * 3);
  if (nrm3) s->nrm.assign(nrm3, nrm3 + (size_t)nv * 3); else s->nrm.assign((size_t)nv * 3, 0.f);
  s->mats.resize(nmat);
  if (nmat) std::memcpy(s->mats.data(), materials, sizeof(Material) * nmat);
  std::vector<u32> order;
  if (!build_scene_bvh(*s, pos3, idx3, max_depth, order)) { delete s; return nullptr; }
  s->tri.resize((size_t)nt * 3); s->tmat.resize(nt); s->orig.resize(nt);
  s->p0.resize((size_t)nt * 3); s->e0.resize((size_t)nt * 3); s->e1.resize((size_t)nt * 3);
  for (u32 i = 0; i < nt; i++) {
    u32 o = order[i];
    s->orig[i] = o;
    for (int k = 0; k < 3; k++) s->tri[i * 3 + k] = idx3[o * 3 + k];
    s->tmat[i] = mat ? mat[o] : 0;
    const float* a = &pos3[idx3[o * 3] * 3];
    const float* b = &pos3[idx3[o * 3 + 1] * 3];
    const float* c = &pos3[idx3[o * 3 + 2] * 3];
    for (int k = 0; k < 3; k++) { s->p0[i * 3 + k] = a[k]; s->e0[i * 3 + k] = b[k] - a[k]; s->e1[i * 3 + k] = c[k] - a[k]; }
  }
  // updateLightTriCudaMem (cudaRTBDPTStream.cu:57-77): host float length of emissive
  for (u32 i = 0; i < nt; i++) {
    const float* e = s->mats.empty() ? nullptr : s->mats[s->tmat[i]].emissive;
    if (e && std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) > 0.f) s->lightTri.push_back(i);
  }
  return s;
}
void orc_scene_destroy(void* h) { delete (Scene*)h; }
// Environment light: W x H RGB texels (W = 0 removes it).
void orc_scene_set_env(void* h, u32 W, u32 H, const float* rgb) {
  Scene* s = (Scene*)h;
  s->envW = W; s->envH = H;
  s->env.assign(rgb, rgb + (size_t)W * H * 3 * (W && H ? 1 : 0));
  if (!W || !H) { s->envW = s->envH = 0; s->env.clear(); }
}
void orc_env_radiance(void* h, const double* d, int prec, double* out) {
  const Scene* s = (const Scene*)h;
  if (prec == 64) put3(env_radiance<double>(*s, v3d<double>(d)), out);
  else put3(env_radiance<float>(*s, v3d<float>(d)), out);
}
// Per-vertex texture coordinates (2 floats per vertex).
void orc_scene_set_uv(void* h, const float* uv2) {
  Scene* s = (Scene*)h;
  s->uv.assign(uv2, uv2 + (size_t)s->nv * 2);
}
// Appends a w x h RGBA texture; returns its id.
u32 orc_scene_add_texture(void* h, u32 w, u32 hgt, const float* rgba) {
  Scene* s = (Scene*)h;
  const u32 first = (u32)(s->tex.size() / 4);
  s->tex.insert(s->tex.end(), rgba, rgba + (size_t)w * hgt * 4);
  s->texDesc.push_back(first);
  s->texDesc.push_back(w);
  s->texDesc.push_back(hgt);
  return (u32)(s->texDesc.size() / 3 - 1);
}
// One texture lookup (the tex_sample above), FP32 or FP64.
// One BRDF sample (FP32) at a hit: hs = norm xyz, diff xyz, trans, specular,
// metallic, roughness; the RNG is PCG32 (state, sequence).  out = new origin,
// direction, throughput (9 floats); returns the new ray type.
int orc_sample_bounce(const float* ro, const float* rd, float t, const float* hs, const float* nl, u64 state, u64 seq,
                      int tir, float* out) {
  HitShade<float> h;
  h.norm = mk(hs[0], hs[1], hs[2]);
  h.pos = mk(0.f, 0.f, 0.f);
  h.diff = mk(hs[3], hs[4], hs[5]);
  h.emissive = mk(0.f, 0.f, 0.f);
  h.trans = hs[6]; h.specular = hs[7]; h.metallic = hs[8]; h.roughness = hs[9];
  Pcg32 rng;
  rng.state = state;
  rng.inc = (seq << 1u) | 1u;
  Bounce<float> b = sample_bounce<float>(rng, mk(ro[0], ro[1], ro[2]), mk(rd[0], rd[1], rd[2]), t, h,
                                         mk(nl[0], nl[1], nl[2]), RT_EYE, false, tir != 0);
  out[0] = b.orig.x; out[1] = b.orig.y; out[2] = b.orig.z;
  out[3] = b.dir.x; out[4] = b.dir.y; out[5] = b.dir.z;
  out[6] = b.mul.x; out[7] = b.mul.y; out[8] = b.mul.z;
  return b.type;
}

// Shading building blocks one at a time (cudaRTCommon.cu:284-345,
// cudaRTBDPTStream.cu:642-663), in FP32 (prec 32) or FP64, for the CPU tests
// that pin them by quadrature, closed forms and sampling statistics.

double orc_d_ggx(int prec, double rough, double noh) {
  return prec == 64 ? d_ggx<double>(rough, noh) : (double)d_ggx<float>((float)rough, (float)noh);
}
double orc_vis_smith(int prec, double rough, double nov, double nol) {
  return prec == 64 ? vis_smith<double>(rough, nov, nol) : (double)vis_smith<float>((float)rough, (float)nov, (float)nol);
}
void orc_sincos_def(const float* x, uint32_t n, float* s, float* c) {
  for (uint32_t i = 0; i < n; i++) sincos_f(x[i], s[i], c[i]);
}
void orc_atan2_def(const float* y, const float* x, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; i++) out[i] = atan2_f(y[i], x[i]);
}
void orc_acos_def(const float* y, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; i++) out[i] = acos_r<float>(y[i]);
}
void orc_f_schlick(int prec, const double* f0, double voh, double* out) {
  if (prec == 64) put3(f_schlick<double>(v3d<double>(f0), voh), out);
  else put3(f_schlick<float>(v3d<float>(f0), (float)voh), out);
}
void orc_sample_ggx(int prec, double xi0, double xi1, double rough, const double* n, double* out) {
  if (prec == 64) put3(sample_ggx<double>(xi0, xi1, rough, v3d<double>(n)), out);
  else put3(sample_ggx<float>((float)xi0, (float)xi1, (float)rough, v3d<float>(n)), out);
}
void orc_basis(int prec, const double* w, double* u, double* v) {
  if (prec == 64) { V3<double> a, b; basis(v3d<double>(w), a, b); put3(a, u); put3(b, v); }
  else { V3<float> a, b; basis(v3d<float>(w), a, b); put3(a, u); put3(b, v); }
}
// a = outDir3 inDir3 irrad3 n3 diff3 metallic roughness specular diffW specW
void orc_shading(int prec, const double* a, double* out) {
  if (prec == 64)
    put3(shading<double>(v3d<double>(a), v3d<double>(a + 3), v3d<double>(a + 6), v3d<double>(a + 9),
                         v3d<double>(a + 12), a[15], a[16], a[17], a[18], a[19]), out);
  else
    put3(shading<float>(v3d<float>(a), v3d<float>(a + 3), v3d<float>(a + 6), v3d<float>(a + 9), v3d<float>(a + 12),
                        (float)a[15], (float)a[16], (float)a[17], (float)a[18], (float)a[19]), out);
}
// n independent BRDF samples of the eye bounce at one hit, sample i drawing
// from PCG32 stream i with the state hashed from (seed, i).  hs = norm3 diff3 trans specular metallic
// roughness.  Per sample: the branch (RT_SPEC 2 / RT_DIFF 1), the next ray's
// origin and direction, and its throughput weight.
void orc_sample_bounce_many(int prec, const double* ro, const double* rd, double t, const double* hs,
                            const double* nl, u64 seed, u32 n, int refl_offset, i32* type, double* orig,
                            double* dir, double* mul) {
  if (prec == 64) bounce_many<double>(ro, rd, t, hs, nl, seed, n, refl_offset, type, orig, dir, mul);
  else bounce_many<float>(ro, rd, t, hs, nl, seed, n, refl_offset, type, orig, dir, mul);
}

// Disney BRDF checks.  m = base3 metallic subsurface specular roughness sheen
// sheenTint clearcoat clearcoatGloss (11 values).
static Disney<double> disney_from(const double* m) {
  Disney<double> d;
  d.base = mk(m[0], m[1], m[2]); d.metallic = m[3]; d.subsurface = m[4]; d.specular = m[5]; d.roughness = m[6];
  d.sheen = m[7]; d.sheenTint = m[8]; d.clearcoat = m[9]; d.clearcoatGloss = m[10];
  return d;
}
void orc_disney_f(const double* m, const double* n, const double* wo, const double* wi, double* out) {
  put3(disney_f<double>(disney_from(m), v3d<double>(n), v3d<double>(wo), v3d<double>(wi)), out);
}
double orc_disney_pdf(const double* m, const double* n, const double* wo, const double* wi) {
  return disney_pdf<double>(disney_from(m), v3d<double>(n), v3d<double>(wo), v3d<double>(wi));
}
// count samples (independent hashed streams from seed): direction and valid flag
void orc_disney_sample_many(const double* m, const double* n, const double* wo, u64 seed, u32 count, double* wi,
                            i32* valid) {
  const Disney<double> d = disney_from(m);
  for (u32 i = 0; i < count; i++) {
    Pcg32 rng; pcg32_seed(rng, splitmix64_mix(seed + (u64)i * 0x9E3779B97F4A7C15ull), i);
    double u0 = uniform<double>(rng), u1 = uniform<double>(rng), u2 = uniform<double>(rng);
    V3<double> w;
    valid[i] = disney_sample<double>(d, v3d<double>(n), v3d<double>(wo), u0, u1, u2, w) ? 1 : 0;
    put3(w, wi + (size_t)i * 3);
  }
}

void orc_tex_sample(void* h, u32 id, double u, double v, int bits, double* out) {
  const Scene* s = (const Scene*)h;
  if (bits == 64) {
    double c[4];
    tex_sample<double>(*s, id, u, v, c);
    for (int k = 0; k < 4; k++) out[k] = c[k];
  } else {
    float c[4];
    tex_sample<float>(*s, id, (float)u, (float)v, c);
    for (int k = 0; k < 4; k++) out[k] = c[k];
  }
}

// Exports the compact BVH: bounds6 (min xyz, max xyz), offOrTSTN2, reorder, nodeN.
u32 orc_scene_bvh(void* h, float* bounds6, u32* off2, u32* reorder, u32 cap) {
  Scene* s = (Scene*)h;
  if (bounds6 && cap >= s->nodeN) {
    for (u32 i = 0; i < s->nodeN; i++) {
      for (int k = 0; k < 3; k++) { bounds6[i * 6 + k] = s->bmin[i * 3 + k]; bounds6[i * 6 + 3 + k] = s->bmax[i * 3 + k]; }
      off2[i * 2] = s->offOrTSTN[i * 2]; off2[i * 2 + 1] = s->offOrTSTN[i * 2 + 1];
    }
  }
  if (reorder) for (u32 i = 0; i < s->nt; i++) reorder[i] = s->orig[i];
  return s->nodeN;
}
u32 orc_scene_light_tris(void* h, u32* out, u32 cap) {
  Scene* s = (Scene*)h;
  for (u32 i = 0; i < s->lightTri.size() && i < cap; i++) out[i] = s->lightTri[i];
  return (u32)s->lightTri.size();
}

// rays: n x 8 floats {ox,oy,oz,maxDist, dx,dy,dz,eps}; hits: n x {u32 tri, f32 t,u,v}.
// mode 0 = brute, 1 = defined (conservative BVH), 2 = reference TracePrimitive.
// prec 32 or 64. stats[2] = stack drops, cap hits (mode 2).
// The BVH debug renderer's depth queries (FP32), rays as for orc_trace (max
// distance r[3]): spec = 0xFFFFFFFF: TraceDepth, out[3i..] = {hit, depth, leaf};
// otherwise TraceDepthParent(spec): {hit, depth (0xFFFFFFFF for -1), parent}.
void orc_trace_depth(void* h, const float* rays, u32 n, u32 spec, u32* out) {
  Scene* s = (Scene*)h;
#pragma omp parallel for schedule(dynamic, 256)
  for (i32 i = 0; i < (i32)n; i++) {
    const float* r = rays + (size_t)i * 8;
    V3<float> org = cvt<float>(r), dir = cvt<float>(r + 4);
    u32* q = out + (size_t)i * 3;
    if (spec == 0xFFFFFFFFu) {
      u32 depth = 0; bool leaf = false;
      bool hit = trace_depth_literal<float>(*s, org, dir, r[3], depth, leaf);
      q[0] = hit; q[1] = depth; q[2] = leaf;
    } else {
      i32 depth = -1; u32 parent = 0;
      bool hit = trace_depth_parent_literal<float>(*s, org, dir, r[3], spec, depth, parent);
      q[0] = hit; q[1] = (u32)depth; q[2] = parent;
    }
  }
}

void orc_trace(void* h, const float* rays, u32 n, u32 cull, u32 mode, u32 prec, void* hits, u32* stats) {
  Scene* s = (Scene*)h;
  Hit* out = (Hit*)hits;
  u32 drops = 0, caps = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : drops, caps)
  for (i32 i = 0; i < (i32)n; i++) {
    const float* r = rays + (size_t)i * 8;
    RefStats st = {0, 0};
    Hit o;
    if (prec == 64) {
      HitR<double> hh; V3<double> org = cvt<double>(r), dir = cvt<double>(r + 4);
      bool ok = mode == 0 ? trace_brute<double>(*s, org, dir, r[3], r[7], cull != 0, hh)
              : mode == 1 ? trace_defined<double>(*s, org, dir, r[3], r[7], cull != 0, hh)
                          : trace_refbin<double>(*s, org, dir, r[3], r[7], cull != 0, hh, st);
      o.tri = ok ? hh.orig : 0xFFFFFFFFu; o.t = ok ? (float)hh.t : (float)kInf; o.u = ok ? (float)hh.u : 0; o.v = ok ? (float)hh.v : 0;
    } else {
      HitR<float> hh; V3<float> org = cvt<float>(r), dir = cvt<float>(r + 4);
      bool ok = mode == 0 ? trace_brute<float>(*s, org, dir, r[3], r[7], cull != 0, hh)
              : mode == 1 ? trace_defined<float>(*s, org, dir, r[3], r[7], cull != 0, hh)
                          : trace_refbin<float>(*s, org, dir, r[3], r[7], cull != 0, hh, st);
      o.tri = ok ? hh.orig : 0xFFFFFFFFu; o.t = ok ? hh.t : (float)kInf; o.u = ok ? hh.u : 0; o.v = ok ? hh.v : 0;
    }
    out[i] = o;
    drops += st.stack_drops; caps += st.cap_hits;
  }
  if (stats) { stats[0] = drops; stats[1] = caps; }
}

// Stateful renderer: the running mean persists across calls like the
// reference's namespace globals.  prec 32 or 64.
struct OrcRenderer { RenderParams p; u32 prec; FrameState<float> f32; FrameState<double> f64; };

void* orc_renderer_create(const void* params, u32 prec) {
  OrcRenderer* r = new OrcRenderer();
  std::memcpy(&r->p, params, sizeof(RenderParams));
  r->prec = prec;
  size_t n = (size_t)r->p.width * r->p.height;
  r->f32.result.assign(n * 3, 0.f); r->f32.N.assign(n, 0u);
  r->f64.result.assign(n * 3, 0.0); r->f64.N.assign(n, 0u);
  return r;
}
void orc_renderer_destroy(void* h) { delete (OrcRenderer*)h; }
u32 orc_render_params_size() { return (u32)sizeof(RenderParams); }

// Renders frames [frame0, frame0+n). result (W*H*3, float) and N (W*H) out;
// gbuf (W*H*9 float, optional) receives the first-hit pos/norm/diff of the last frame.
void orc_render(void* scene, void* rh, u32 frame0, u32 nframes, float* result, u32* N, float* gbuf, u32* lvN) {
  Scene* s = (Scene*)scene;
  OrcRenderer* r = (OrcRenderer*)rh;
  size_t n = (size_t)r->p.width * r->p.height;
  if (r->prec == 64) {
    std::vector<double> g; if (gbuf) g.assign(n * 9, 0.0);
    if (r->p.mode == 8) render_bdpt_regen<double>(*s, r->p, frame0, nframes, r->f64);
    else if (r->p.mode == 7) render_disney<double>(*s, r->p, frame0, nframes, r->f64);
    else if (r->p.mode == 6) render_pt_literal<double>(*s, r->p, frame0, nframes, r->f64);
    else if (r->p.mode >= 4) render_pt<double>(*s, r->p, frame0, nframes, r->f64);
    else if (r->p.mode != 0) render_adaptive<double>(*s, r->p, frame0, nframes, r->f64);
    else render_frames<double>(*s, r->p, frame0, nframes, r->f64, gbuf ? g.data() : nullptr);
    if (result) for (size_t i = 0; i < n * 3; i++) result[i] = (float)r->f64.result[i];
    if (N) std::memcpy(N, r->f64.N.data(), n * 4);
    if (gbuf) for (size_t i = 0; i < n * 9; i++) gbuf[i] = (float)g[i];
    if (lvN) *lvN = r->f64.lvN;
  } else {
    std::vector<float> g; if (gbuf) g.assign(n * 9, 0.f);
    if (r->p.mode == 8) render_bdpt_regen<float>(*s, r->p, frame0, nframes, r->f32);
    else if (r->p.mode == 7) render_disney<float>(*s, r->p, frame0, nframes, r->f32);
    else if (r->p.mode == 6) render_pt_literal<float>(*s, r->p, frame0, nframes, r->f32);
    else if (r->p.mode >= 4) render_pt<float>(*s, r->p, frame0, nframes, r->f32);
    else if (r->p.mode != 0) render_adaptive<float>(*s, r->p, frame0, nframes, r->f32);
    else render_frames<float>(*s, r->p, frame0, nframes, r->f32, gbuf ? g.data() : nullptr);
    if (result) std::memcpy(result, r->f32.result.data(), n * 12);
    if (N) std::memcpy(N, r->f32.N.data(), n * 4);
    if (gbuf) std::memcpy(gbuf, g.data(), n * 36);
    if (lvN) *lvN = r->f32.lvN;
  }
}

// Adaptive-mode state: which = 0 G-buffer pos, 1 norm, 2 diff (W*H*3 floats),
// 3 filtered (W*H*3), 4 error map (W*H), 5 per-pixel sample counts (W*H u32).
// Returns the error sum of the last adaptive frame through *sum when non-null.
void orc_adaptive_state(void* rh, u32 which, void* out, float* sum) {
  OrcRenderer* r = (OrcRenderer*)rh;
  auto dump = [&](auto& fs) {
    size_t n = (size_t)r->p.width * r->p.height;
    float* o = (float*)out;
    auto cp = [&](auto& v) { for (size_t i = 0; i < v.size(); i++) o[i] = (float)v[i]; };
    if (which == 0) cp(fs.gpos);
    else if (which == 1) cp(fs.gnorm);
    else if (which == 2) cp(fs.gdiff);
    else if (which == 3) { if (fs.filtered.size()) std::memcpy(out, fs.filtered.data(), n * 12); }
    else if (which == 4) { if (fs.err.size()) std::memcpy(out, fs.err.data(), n * 4); }
    else if (which == 5) { if (fs.counts.size()) std::memcpy(out, fs.counts.data(), n * 4); }
    else if (which == 6) cp(fs.sq);
    else if (which == 7) *(u32*)out = fs.rounds;
    else if (which == 8) { if (fs.paths.size()) std::memcpy(out, fs.paths.data(), n * 4); }
    if (sum) *sum = fs.err_sum;
  };
  if (r->prec == 64) dump(r->f64); else dump(r->f32);
}

// Defined cross-bilateral (FP32, one exponential per tap) for tests.
void orc_bilateral_defined(const float* color, const float* pos, const float* norm, const float* emit, u32 W, u32 H,
                           float sc, float sp, float sn, float se, u32 radius, float* out) {
  std::vector<float> g; gaussian_table(sc, radius, g);
  bilateral_defined(color, pos, norm, emit, W, H, sc, sp, sn, se, radius, g, out);
}
// The filter with the extension terms: spatial sigma ss (0: the colour sigma)
// and the relative-depth term sz (0: off) about camera position cam.
// prec 32: the defined one-exponential arithmetic; 64: the literal product.
void orc_bilateral_ext(const float* color, const float* pos, const float* norm, const float* emit, u32 W, u32 H,
                       float sc, float sp, float sn, float se, float ss, float sz, const float* cam, u32 radius,
                       u32 prec, float* out) {
  std::vector<float> g; gaussian_table(ss > 0.f ? ss : sc, radius, g);
  if (prec == 64) bilateral<double>(color, pos, norm, emit, W, H, sc, sp, sn, se, radius, g, out, sz, cam);
  else bilateral_defined(color, pos, norm, emit, W, H, sc, sp, sn, se, radius, g, out, sz, cam);
}
float orc_expf(float x) { return cpt_expf(x); }
double orc_tree_sum(const float* e, u32 n) { return tree_sum(e, n); }

// Double-precision result of the running mean (for the FP64 gate).
void orc_render_counts(void* rh, u64* out) {
  OrcRenderer* r = (OrcRenderer*)rh;
  out[0] = r->prec == 64 ? r->f64.ext_rays : r->f32.ext_rays;
  out[1] = r->prec == 64 ? r->f64.shadow_rays : r->f32.shadow_rays;
}
int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_render_result64(void* rh, double* out) {
  OrcRenderer* r = (OrcRenderer*)rh;
  std::memcpy(out, r->f64.result.data(), r->f64.result.size() * 8);
}

// Light-vertex cache of the last light pass: n x 21 floats
// {pos3, norm3, irrad3, irradDir3, diff3, emissive3, spec, metal, rough}.
u32 orc_light_vertices(void* rh, float* out, u32 cap) {
  OrcRenderer* r = (OrcRenderer*)rh;
  auto dump = [&](auto& fs) {
    for (u32 i = 0; i < fs.lvN && i < cap; i++) {
      auto& v = fs.lv[i];
      float* o = out + (size_t)i * 21;
      auto put = [&](int k, auto a) { o[k] = (float)a.x; o[k + 1] = (float)a.y; o[k + 2] = (float)a.z; };
      put(0, v.pos); put(3, v.norm); put(6, v.irrad); put(9, v.irradDir); put(12, v.diff); put(15, v.emissive);
      o[18] = (float)v.specular; o[19] = (float)v.metallic; o[20] = (float)v.roughness;
    }
    return fs.lvN;
  };
  return r->prec == 64 ? dump(r->f64) : dump(r->f32);
}

void orc_bilateral(const float* color, const float* pos, const float* norm, const float* emit, u32 W, u32 H,
                   float sc, float sp, float sn, float se, u32 radius, u32 prec, float* out) {
  std::vector<float> g; gaussian_table(sc, radius, g);
  if (prec == 64) bilateral<double>(color, pos, norm, emit, W, H, sc, sp, sn, se, radius, g, out);
  else bilateral<float>(color, pos, norm, emit, W, H, sc, sp, sn, se, radius, g, out);
}
void orc_gaussian_table(float sigma, u32 radius, float* out) {
  std::vector<float> g; gaussian_table(sigma, radius, g);
  std::memcpy(out, g.data(), g.size() * 4);
}

// pt_calculateSquareError_kernel (AdapBilatAtomic.cu:1143-1156)
void orc_square_error(const float* filtered, const float* raw, const u32* N, u32 nd, u32 n, float* out) {
  for (u32 x = 0; x < n; x++) {
    u32 sn = N[x];
    float red = (float)nd / (float)(nd + sn);
    if (nd + sn < sn) red = 0;
    float d0 = filtered[x * 3] - raw[x * 3], d1 = filtered[x * 3 + 1] - raw[x * 3 + 1], d2 = filtered[x * 3 + 2] - raw[x * 3 + 2];
    out[x] = red * ((d0 * d0 + d1 * d1) + d2 * d2) / 3.f;
  }
}
// pt_calculateEmpiricalSquareError_kernel (AdapEmp.cu:949-960)
void orc_empirical_error(const float* mean, const float* sq, const u32* N, u32 n, float* out) {
  for (u32 x = 0; x < n; x++) {
    u32 sn = N[x];
    float e = ((sq[x * 3] - mean[x * 3] * mean[x * 3]) + (sq[x * 3 + 1] - mean[x * 3 + 1] * mean[x * 3 + 1])) +
              (sq[x * 3 + 2] - mean[x * 3 + 2] * mean[x * 3 + 2]);
    e = e * (float)sn / ((float)(sn - 1) * 3.f);
    out[x] = std::fmin(e, 1.0f);
  }
}
// Round-based allocation of a given error map (alloc_rounds); returns the rounds run.
u32 orc_alloc_rounds(const float* err, u32 n, u32 gen, u32 bisect, float desired, float min_prob, u64 seed,
                     u32 frame, u32* counts) {
  RenderParams p;
  std::memset(&p, 0, sizeof p);
  p.seed = seed;
  std::vector<float> e(err, err + n);
  std::vector<u32> c;
  u32 rounds = 0;
  alloc_rounds(p, frame, e, n, gen, bisect != 0, desired, min_prob, c, rounds);
  std::memcpy(counts, c.data(), sizeof(u32) * n);
  return rounds;
}

// pt_genAdapPathQueue_kernel counts (AdapBilatAtomic.cu:830-831): floor(gen * e/sum)
void orc_alloc_proportional(const float* err, u32 n, float sum, u32 gen, u32* counts) {
  for (u32 x = 0; x < n; x++) { float r = err[x] / sum; counts[x] = (u32)((float)gen * r); }
}

}  // extern "C"
