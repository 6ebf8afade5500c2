// Device float arithmetic with the operation sequence of the reference's
// cudahelper.h (include/cudahelper.h:119-304): explicit fma nesting in dot and
// cross, correctly rounded scalar divide, round-down sqrt in length(); a vector
// over a scalar multiplies by the correctly rounded reciprocal.  The library is
// compiled with -ffp-contract=off, so every other a*b+c stays two roundings.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpt {

struct f3 { float x, y, z; };

__device__ __forceinline__ f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
// A vector over a scalar: one correctly rounded reciprocal, then three
// products (within an ulp of the reference's three divisions; the FP32 oracle
// evaluates it the same way).
__device__ __forceinline__ f3 operator/(f3 a, float s) { const float r = 1.f / s; return mk3(a.x * r, a.y * r, a.z * r); }
__device__ __forceinline__ f3 vmul(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 vmax0(f3 a) { return mk3(fmaxf(0.f, a.x), fmaxf(0.f, a.y), fmaxf(0.f, a.z)); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return mk3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
// __fsqrt_rd realised as a correctly rounded sqrt stepped down when it rounded up.
__device__ __forceinline__ float sqrt_rd(float x) {
  float s = sqrtf(x);
  if (s > 0.f && fmaf(s, s, -x) > 0.f) s = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, s) - 1u);
  return s;
}
__device__ __forceinline__ float length(f3 a) { return sqrt_rd(fmaf(a.z, a.z, fmaf(a.y, a.y, a.x * a.x))); }
__device__ __forceinline__ f3 normalize(f3 a) { return a / length(a); }
__device__ __forceinline__ float lerpf(float x, float y, float d) { return x + (y - x) * d; }
__device__ __forceinline__ f3 vlerp(f3 x, f3 y, float d) { return x + (y - x) * d; }
__device__ __forceinline__ float saturatef_(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }
__device__ __forceinline__ bool isinf3(f3 a) { return isinf(a.x) || isinf(a.y) || isinf(a.z); }
__device__ __forceinline__ f3 ld3(const float4 v) { return mk3(v.x, v.y, v.z); }

// ---- trigonometry with one defined operation sequence ---------------------
// The host oracle's FP32 restatement evaluates these same float operations,
// so the two agree bit for bit where two libm's sinf / atan2f would differ
// in the last place now and then.
//
// sin and cos of x >= 0 (the sampling angles 2 pi u, u in (0, 1]): the
// octant reduction of Cephes' sinf / cosf (S. L. Moshier, Cephes Math Library,
// 1989): j = trunc(x 4/pi) rounded up to even, r = x - j pi/4 with pi/4 in three
// parts (each product applied with an fma), then Cephes' minimax polynomials
// for sin and cos on [-pi/4, pi/4]; quadrant j/2 mod 4 picks and signs them.
__device__ __forceinline__ void sincos_def(float x, float& s, float& c) {
  int j = (int)(x * 1.27323954473516f);
  j += j & 1;
  const float y = (float)j;
  float r = fmaf(-y, 0.78515625f, x);
  r = fmaf(-y, 2.4187564849853515625e-4f, r);
  r = fmaf(-y, 3.77489497744594108e-8f, r);
  const float z = r * r;
  const float ps = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f), z * r, r);
  const float pc = fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f),
                        z * z, fmaf(-0.5f, z, 1.f));
  const int q = (j >> 1) & 3;
  s = (q & 1) ? pc : ps;
  c = (q & 1) ? ps : pc;
  if (q >= 2) s = -s;
  if (q == 1 || q == 2) c = -c;
}
// atan of t in [0, 1]: Cephes' atanf (Moshier), t above tan(pi/8) moved to
// (t - 1) / (t + 1) around pi/4, then its minimax polynomial.
__device__ __forceinline__ float atan01_def(float t) {
  float y0 = 0.f;
  if (t > 0.4142135623730950f) {
    y0 = 0.78539816339744831f;
    t = (t - 1.f) / (t + 1.f);
  }
  const float z = t * t;
  const float p = fmaf(fmaf(fmaf(8.05374449538e-2f, z, -1.38776856032e-1f), z, 1.99777106478e-1f), z,
                       -3.33329491539e-1f);
  return y0 + fmaf(p * z, t, t);
}
// atan2(y, x) in [-pi, pi] from atan01_def by octant.
__device__ __forceinline__ float atan2_def(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  float a;
  if (ay <= ax) a = ax > 0.f ? atan01_def(ay / ax) : 0.f;
  else a = 1.57079632679489662f - atan01_def(ax / ay);
  if (x < 0.f) a = 3.14159265358979324f - a;
  return y < 0.f ? -a : a;
}
// acos(y) for y in [-1, 1] as atan2(sqrt((1 - y)(1 + y)), y).
__device__ __forceinline__ float acos_def(float y) { return atan2_def(sqrtf((1.f - y) * (1.f + y)), y); }

// ---- PCG32 (O'Neill 2014), uniforms in (0, 1] with 24 bits -------------
__device__ __forceinline__ uint32_t pcg32_next(uint64_t& state, uint64_t inc) {
  uint64_t old = state;
  state = old * 6364136223846793005ULL + inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((-rot) & 31u));
}
__device__ __forceinline__ void pcg32_seed(uint64_t& state, uint64_t inc, uint64_t initstate) {
  state = 0u;
  pcg32_next(state, inc);
  state += initstate;
  pcg32_next(state, inc);
}
__device__ __forceinline__ float uniform01(uint64_t& state, uint64_t inc) {
  return (float)((pcg32_next(state, inc) >> 8) + 1u) * (1.0f / 16777216.0f);
}

}  // namespace cpt
