// Device float arithmetic with the operation sequence of the reference's
// cudahelper.h (include/cudahelper.h:119-304): explicit fma nesting in dot and
// cross, correctly rounded divide, round-down sqrt in length().  The library is
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
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
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
__device__ __forceinline__ f3 normalize(f3 a) { float l = length(a); return mk3(a.x / l, a.y / l, a.z / l); }
__device__ __forceinline__ float lerpf(float x, float y, float d) { return x + (y - x) * d; }
__device__ __forceinline__ f3 vlerp(f3 x, f3 y, float d) { return x + (y - x) * d; }
__device__ __forceinline__ float saturatef_(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }
__device__ __forceinline__ bool isinf3(f3 a) { return isinf(a.x) || isinf(a.y) || isinf(a.z); }
__device__ __forceinline__ f3 ld3(const float4 v) { return mk3(v.x, v.y, v.z); }

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
