// Per-pixel error MLP 9 -> 64 -> 64 -> 1 (GELU), BF16 operands on gfx950 MFMA
// (v_mfma_f32_16x16x32_bf16, and v_mfma_f32_16x16x16_bf16 for the 9-feature
// layer) with FP32 accumulation, FP32 master weights and Adam.  No reference
// counterpart: CudaPT links FANN but never calls it; the network replaces the
// filter-residual error of the adaptive renderer
// (src/cudaRTBDPTStreamAdapBilatAtomic.cu:1143-1156) as the allocation signal.
//
// Training: a workgroup of four waves; each wave takes 64 pixels at a time
// and keeps the layer activations in registers between layers (the k-permuted
// operand scheme below), accumulating its weight gradients over its chunks in
// MFMA accumulators:
//   forward   Z1 = X W1 + b1, H1 = gelu(Z1), Z2 = H1 W2 + b2, H2 = gelu(Z2), y = H2 w3 + b3
//   backward  dy = 2 (y - t) / B, dZ2 = dy w3^T * gelu'(Z2), dH1 = dZ2 W2^T,
//             dZ1 = dH1 * gelu'(Z1), dW2 = H1^T dZ2, dW1 = X^T dZ1 (all MFMA)
// The waves' gradients are summed in wave order into one partial per
// workgroup; k_mlp_adam sums the partials in a fixed order (deterministic) and
// applies Adam (Kingma & Ba 2015).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.h"

namespace cpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 with round-to-nearest-even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form (Hendrycks & Gimpel 2016): 0.5 x (1 + tanh u) with
// u = c (x + a x^3), evaluated as x sigmoid(2u) = x / (1 + 2^(-2u log2 e)) with
// the hardware exponential and reciprocal (relative error ~1e-7, far below the
// BF16 rounding of the activations).  One exponential, one reciprocal and four
// other operations; through tanh with a division (correctly rounded under this
// build's flags) inference took 248 us per 1920 x 1080 pass against 116 us.
// Large |x| saturate to x and -0.
__device__ __forceinline__ float gelu_sigmoid(float x, float x2) {
  const float k = -2.f * 0.7978845608028654f * 1.4426950408889634f;   // -2 c log2(e)
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * __builtin_fmaf(k * 0.044715f, x2, k)));
}
__device__ __forceinline__ float gelu_val(float x) { return x * gelu_sigmoid(x, x * x); }
// With the derivative, for training: with s = sigmoid(2u),
// d/dx [x s] = s + x s (1 - s) 2c (1 + 3 a x^2).
__device__ __forceinline__ void gelu_fwd(float x, float& g, float& dg) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float x2 = x * x;
  const float sg = gelu_sigmoid(x, x2);
  g = x * sg;
  dg = __builtin_fmaf(g * (1.f - sg), (2.f * c) * __builtin_fmaf(3.f * a, x2, 1.f), sg);
}

// ---- register-resident layers ----------------------------------------------
// A wave computes its 64 pixels transposed: activations are [unit][pixel]
// 64 x 64 matrices held as 4 x 4 tiles of 16 x 16 in the MFMA output layout,
// lane l = 16 g + c holding rows 4g..4g+3 of column c of every tile.  The next
// layer's MFMA reads them as its B operand without leaving the registers: the
// contraction index k of v_mfma_f32_16x16x32_bf16 may be permuted freely if
// A and B agree, so k-slot j of lane group g in k-block kb stands for unit
//   perm(kb, g, j) = (2 kb + j / 4) * 16 + 4 g + j % 4,
// which is exactly what the lane holds of tiles 2kb and 2kb+1.  The weight
// operands are packed on the device in that order (mlp_pack_one).  Only the
// weight gradients contract over pixels and need the pixel index in the
// k-slots: their operands go through LDS, one 8 KB [unit][pixel] BF16 plane
// each.  Layer 1 (K = 9 features) uses v_mfma_f32_16x16x16_bf16 with K padded
// to 16 (the 16x16x32 form would pad it to 32).
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 ld_frag8(const uint16_t* p) { return *(const bf16x8*)p; }
__device__ __forceinline__ s16x4 ld_frag4(const uint16_t* p) { return *(const s16x4*)p; }
__device__ __forceinline__ __bf16 as_bf(uint16_t h) { return __builtin_bit_cast(__bf16, h); }

// LDS order within a wave: LDS instructions of one wave execute in issue
// order; this keeps the compiler from moving them across the point and waits
// for the outstanding ones.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

constexpr int kTrainWaves = 4;                  // waves per training workgroup
// Training takes 32 pixels per wave at a time (two pixel blocks): the
// activations of a chunk then fit the register file next to the
// weight-gradient accumulators; the weight gradients contract over the
// chunk's 32 pixels, one k-block of v_mfma_f32_16x16x32_bf16.
constexpr int kTrainPB = 2;                      // 16-pixel blocks per training chunk
constexpr int kTrainChunk = 16 * kTrainPB;
// Row stride of the transposed planes: 8 pads a 64-byte row to 80 bytes, so
// the four lane groups' 16-bit stores of one unit quad and the 16 rows of a
// 16-byte fragment read fall in different LDS banks (with 64-byte rows both
// hit one bank four times over).
constexpr int kPlaneStride = kTrainChunk + 8;
struct MlpWaveLds {
  alignas(16) uint16_t xt[16 * kPlaneStride];    // X^T  [feature][pixel] (features 9..15 zero)
  alignas(16) uint16_t h1t[64 * kPlaneStride];   // H1^T [hidden][pixel], then dZ1^T
  alignas(16) uint16_t dz2t[64 * kPlaneStride];  // dZ2^T [out][pixel]
  float vec[3][64];                             // dW3, db2, db1 summed over the wave's chunks
};
// Weight operands and vectors, shared by the workgroup's waves.
struct MlpWeightLds {
  alignas(16) uint16_t w1f[kMlpW1Frag];
  alignas(16) uint16_t w2tf[kMlpW2Frag];
  alignas(16) uint16_t w2f[kMlpW2Frag];
  float b1[64], b2[64], w3[64];
};
struct MlpTrainLds {
  MlpWeightLds w;
  union {
    MlpWaveLds wave[kTrainWaves];
    float red[kMlpParams + 3];                  // the workgroup's gradient, summed wave by wave
  } u;
};

template <bool TRAIN>
__device__ __forceinline__ void stage_weights(MlpWeightLds& s, const MlpDevWeights& w, int tid, int nthreads) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  for (int i = tid; i < kMlpW1Frag / 8; i += nthreads) ((u32x4*)s.w1f)[i] = ((const u32x4*)w.w1f)[i];
  for (int i = tid; i < kMlpW2Frag / 8; i += nthreads) {
    ((u32x4*)s.w2tf)[i] = ((const u32x4*)w.w2tf)[i];
    if constexpr (TRAIN) ((u32x4*)s.w2f)[i] = ((const u32x4*)w.w2f)[i];
  }
  if (tid < 64) { s.b1[tid] = w.b1[tid]; s.b2[tid] = w.b2[tid]; s.w3[tid] = w.w3[tid]; }
}

// Layer-1 B operands of the wave's pixels: lane (g, c) holds features
// 4g..4g+3 of pixel base + 16 pb + c (zero past feature 8 and past n).
template <int PB>
__device__ __forceinline__ void load_x(s16x4 xb[PB], const float* __restrict__ x, uint32_t base, uint32_t n, int g,
                                       int c) {
#pragma unroll
  for (int pb = 0; pb < PB; pb++) {
    const uint32_t px = base + pb * 16 + c;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int f = g * 4 + j;
      const float v = (f < kMlpIn && px < n) ? x[(size_t)px * kMlpIn + f] : 0.f;
      xb[pb][j] = (short)f2bf(v);
    }
  }
}

// Z1^T tile [hidden block hb][pixel block pb].
__device__ __forceinline__ f32x4 layer1_tile(const uint16_t* w1f, int hb, const s16x4 xb, int l) {
  return mfma16(ld_frag4(w1f + (hb * 64 + l) * 4), xb, f32x4{0.f, 0.f, 0.f, 0.f});
}
// Tile [row block rb][pixel block pb] of A (64 x 64, packed per row block and
// k-block) times the 64 x 64 register matrix given by its k-permuted B
// operands bk[kb][pb].
template <int PB>
__device__ __forceinline__ f32x4 layer64_tile(const uint16_t* a, int rb, int pb, const bf16x8 bk[2][PB], int l) {
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma32(ld_frag8(a + ((rb * 2 + 0) * 64 + l) * 8), bk[0][pb], acc);
  acc = mfma32(ld_frag8(a + ((rb * 2 + 1) * 64 + l) * 8), bk[1][pb], acc);
  return acc;
}

typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
// Keeps the scheduler from hoisting the next tile's loads and MFMAs above the
// current tile's element work: with every tile in flight at once the live
// set exceeds the register file.
__device__ __forceinline__ void tile_fence() { __builtin_amdgcn_sched_barrier(0); }
// Sums v over the 16 pixel lanes of each lane group and adds it to acc[unit]
// (lane c == 0 of each group writes its four units).
__device__ __forceinline__ void add_unit_sums(float* acc, float v[4][4], int g, int c, bool first) {
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      float t = v[b][e];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) t += __shfl_xor(t, o, 64);
      if (c == 0) {
        const int u = b * 16 + g * 4 + e;
        acc[u] = first ? t : acc[u] + t;
      }
    }
}

__global__ __launch_bounds__(64 * kTrainWaves, 1) void k_mlp_train(MlpDevWeights w, const float* __restrict__ x,
                                                                   const float* __restrict__ target, uint32_t n,
                                                                   float inv_batch, float* __restrict__ part,
                                                                   float* __restrict__ loss_part) {
  __shared__ MlpTrainLds s;
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
  stage_weights<true>(s.w, w, tid, 64 * kTrainWaves);
  for (int i = l; i < 3 * 64; i += 64) s.u.wave[wave].vec[i / 64][i % 64] = 0.f;
  __syncthreads();
  const float b3 = w.b3[0];
  MlpWaveLds& L = s.u.wave[wave];
  // weight-gradient accumulators over the wave's chunks (MFMA accumulators)
  f32x4 aw2[4][4], aw1[4];
  float ab3 = 0.f, aloss = 0.f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    aw1[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 4; j++) aw2[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  constexpr int PB = kTrainPB, C = kTrainChunk, S = kPlaneStride;
  const uint32_t nchunks = (n + C - 1) / C;
  for (uint32_t ch = blockIdx.x * kTrainWaves + wave; ch < nchunks; ch += gridDim.x * kTrainWaves) {
    const uint32_t base = ch * C;
    s16x4 xb[PB];
    load_x<PB>(xb, x, base, n, g, c);
#pragma unroll
    for (int pb = 0; pb < PB; pb++)
#pragma unroll
      for (int j = 0; j < 4; j++) L.xt[(g * 4 + j) * S + pb * 16 + c] = (uint16_t)xb[pb][j];
    // ---- forward; the GELU derivatives and H2 are kept as FP16 pairs
    bf16x8 hk[2][PB];                                      // H1^T as layer-2 B operands
    h16x2 dg1[4][PB][2];
#pragma unroll
    for (int hb = 0; hb < 4; hb++)
#pragma unroll
      for (int pb = 0; pb < PB; pb++) {
        const f32x4 z = layer1_tile(s.w.w1f, hb, xb[pb], l);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int u = hb * 16 + g * 4 + e;
          float gv, dg;
          gelu_fwd(z[e] + s.w.b1[u], gv, dg);
          dg1[hb][pb][e >> 1][e & 1] = (_Float16)dg;
          const uint16_t hbits = f2bf(gv);
          hk[hb >> 1][pb][(hb & 1) * 4 + e] = as_bf(hbits);
          L.h1t[u * S + pb * 16 + c] = hbits;
        }
        tile_fence();
      }
    h16x2 h2[4][PB][2], dg2[4][PB][2];
    float ys[PB];
#pragma unroll
    for (int pb = 0; pb < PB; pb++) ys[pb] = 0.f;
#pragma unroll
    for (int ob = 0; ob < 4; ob++)
#pragma unroll
      for (int pb = 0; pb < PB; pb++) {
        const f32x4 z = layer64_tile<PB>(s.w.w2tf, ob, pb, hk, l);  // Z2^T
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int u = ob * 16 + g * 4 + e;
          float gv, dg;
          gelu_fwd(z[e] + s.w.b2[u], gv, dg);
          h2[ob][pb][e >> 1][e & 1] = (_Float16)gv;
          dg2[ob][pb][e >> 1][e & 1] = (_Float16)dg;
          ys[pb] = fmaf(gv, s.w.w3[u], ys[pb]);
        }
        tile_fence();
      }
    float dy[PB];
#pragma unroll
    for (int pb = 0; pb < PB; pb++) {
      float y = ys[pb];
      y += __shfl_xor(y, 16, 64);
      y += __shfl_xor(y, 32, 64);
      y += b3;
      const uint32_t px = base + pb * 16 + c;
      const float d = px < n ? y - target[px] : 0.f;
      if (g == 0) aloss = fmaf(d, d, aloss);
      dy[pb] = 2.f * d * inv_batch;
      if (g == 0) ab3 += dy[pb];
    }
    // ---- backward
    bf16x8 dk[2][PB];                                      // dZ2^T as B operands
    float v3[4][4], v2[4][4];
#pragma unroll
    for (int ob = 0; ob < 4; ob++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int u = ob * 16 + g * 4 + e;
        const float w3u = s.w.w3[u];
        v3[ob][e] = v2[ob][e] = 0.f;
#pragma unroll
        for (int pb = 0; pb < PB; pb++) {
          v3[ob][e] = fmaf((float)h2[ob][pb][e >> 1][e & 1], dy[pb], v3[ob][e]);
          const float dz = dy[pb] * w3u * (float)dg2[ob][pb][e >> 1][e & 1];
          v2[ob][e] += dz;
          const uint16_t bits = f2bf(dz);
          dk[ob >> 1][pb][(ob & 1) * 4 + e] = as_bf(bits);
          L.dz2t[u * S + pb * 16 + c] = bits;
        }
        tile_fence();
      }
    add_unit_sums(L.vec[0], v3, g, c, false);
    add_unit_sums(L.vec[1], v2, g, c, false);
    wave_lds_sync();
    // dW2[in][out] += H1^T . dZ2 over the chunk's pixels
#pragma unroll
    for (int kb = 0; kb < C / 32; kb++) {
      bf16x8 bo[4];
#pragma unroll
      for (int ob = 0; ob < 4; ob++) bo[ob] = ld_frag8(&L.dz2t[(ob * 16 + c) * S + kb * 32 + g * 8]);
#pragma unroll
      for (int ib = 0; ib < 4; ib++) {
        const bf16x8 ai = ld_frag8(&L.h1t[(ib * 16 + c) * S + kb * 32 + g * 8]);
#pragma unroll
        for (int ob = 0; ob < 4; ob++) aw2[ib][ob] = mfma32(ai, bo[ob], aw2[ib][ob]);
      }
    }
    wave_lds_sync();                                       // H1^T reads done: dZ1^T takes its place
    // dH1^T = W2 . dZ2^T; dZ1 = dH1 * gelu'(Z1)
    float v1[4][4];
#pragma unroll
    for (int ib = 0; ib < 4; ib++) {
#pragma unroll
      for (int e = 0; e < 4; e++) v1[ib][e] = 0.f;
#pragma unroll
      for (int pb = 0; pb < PB; pb++) {
        const f32x4 dh = layer64_tile<PB>(s.w.w2f, ib, pb, dk, l);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int u = ib * 16 + g * 4 + e;
          const float dz = dh[e] * (float)dg1[ib][pb][e >> 1][e & 1];
          v1[ib][e] += dz;
          L.h1t[u * S + pb * 16 + c] = f2bf(dz);
        }
        tile_fence();
      }
    }
    add_unit_sums(L.vec[2], v1, g, c, false);
    wave_lds_sync();
    // dW1[feature][hidden] += X^T . dZ1 over the chunk's pixels
#pragma unroll
    for (int kb = 0; kb < C / 32; kb++) {
      const bf16x8 ax = ld_frag8(&L.xt[c * S + kb * 32 + g * 8]);
#pragma unroll
      for (int hb = 0; hb < 4; hb++)
        aw1[hb] = mfma32(ax, ld_frag8(&L.h1t[(hb * 16 + c) * S + kb * 32 + g * 8]), aw1[hb]);
    }
    wave_lds_sync();                                       // the next chunk rewrites the planes
  }
  for (int o = 1; o < 64; o <<= 1) {
    ab3 += __shfl_xor(ab3, o, 64);
    aloss += __shfl_xor(aloss, o, 64);
  }
  // per-unit vectors of the wave into registers before the planes are reused
  const float vw3 = L.vec[0][l], vb2 = L.vec[1][l], vb1 = L.vec[2][l];
  __syncthreads();                                         // every wave's planes are dead: red takes them
  // the workgroup's gradient: waves add theirs in order
  for (int wv = 0; wv < kTrainWaves; wv++) {
    if (wave == wv) {
      auto put = [&](int p, float v) { s.u.red[p] = wv == 0 ? v : s.u.red[p] + v; };
#pragma unroll
      for (int ib = 0; ib < 4; ib++)
#pragma unroll
        for (int ob = 0; ob < 4; ob++)
#pragma unroll
          for (int e = 0; e < 4; e++) put(kMlpOffW2 + (ib * 16 + g * 4 + e) * 64 + ob * 16 + c, aw2[ib][ob][e]);
#pragma unroll
      for (int hb = 0; hb < 4; hb++)
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (g * 4 + e < kMlpIn) put(kMlpOffW1 + (g * 4 + e) * 64 + hb * 16 + c, aw1[hb][e]);
      put(kMlpOffW3 + l, vw3);
      put(kMlpOffB2 + l, vb2);
      put(kMlpOffB1 + l, vb1);
      if (l == 0) {
        put(kMlpOffB3, ab3);
        put(kMlpParams, aloss);
      }
    }
    __syncthreads();
  }
  float* P = part + (size_t)blockIdx.x * kMlpParams;
  for (int p = tid; p < kMlpParams; p += 64 * kTrainWaves) P[p] = s.u.red[p];
  if (tid == 0) loss_part[blockIdx.x] = s.u.red[kMlpParams];
}

// Inference: a wave takes chunks of 16 * kInferBlocks pixels; kInferMinWaves
// waves per SIMD bound its registers (115 of 128, no spills).
constexpr int kInferBlocks = 4, kInferMinWaves = 4;
constexpr uint32_t kInferMaxGroups = 2048;
__global__ __launch_bounds__(64 * kTrainWaves, kInferMinWaves) void k_mlp_infer(MlpDevWeights w,
                                                                               const float* __restrict__ x,
                                                                               uint32_t n, float* __restrict__ y) {
  constexpr int PB = kInferBlocks;
  __shared__ MlpWeightLds s;
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
  stage_weights<false>(s, w, tid, 64 * kTrainWaves);
  __syncthreads();
  const float b3 = w.b3[0];
  const uint32_t nchunks = (n + 16 * PB - 1) / (16 * PB);
  for (uint32_t ch = blockIdx.x * kTrainWaves + wave; ch < nchunks; ch += gridDim.x * kTrainWaves) {
    // the weights are read from LDS again for every chunk: held in registers
    // across chunks they would take about 90 of them and halve the waves
    asm volatile("" ::: "memory");
    const uint32_t base = ch * (16 * PB);
    s16x4 xb[PB];
    load_x<PB>(xb, x, base, n, g, c);
    bf16x8 hk[2][PB];
#pragma unroll
    for (int hb = 0; hb < 4; hb++)
#pragma unroll
      for (int pb = 0; pb < PB; pb++) {
        const f32x4 z = layer1_tile(s.w1f, hb, xb[pb], l);
#pragma unroll
        for (int e = 0; e < 4; e++)
          hk[hb >> 1][pb][(hb & 1) * 4 + e] = as_bf(f2bf(gelu_val(z[e] + s.b1[hb * 16 + g * 4 + e])));
      }
    float ys[PB];
#pragma unroll
    for (int pb = 0; pb < PB; pb++) ys[pb] = 0.f;
#pragma unroll
    for (int ob = 0; ob < 4; ob++)
#pragma unroll
      for (int pb = 0; pb < PB; pb++) {
        const f32x4 z = layer64_tile<PB>(s.w2tf, ob, pb, hk, l);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int u = ob * 16 + g * 4 + e;
          ys[pb] = fmaf(gelu_val(z[e] + s.b2[u]), s.w3[u], ys[pb]);
        }
      }
#pragma unroll
    for (int pb = 0; pb < PB; pb++) {
      float v = ys[pb];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      const uint32_t px = base + pb * 16 + c;
      if (g == 0 && px < n) y[px] = v + b3;
    }
  }
}

// Writes parameter p (value w) into the operand pack: the k-permuted
// fragment order the register-resident layers read (see above).
__device__ __forceinline__ void mlp_pack_one(MlpDevWeights d, uint32_t p, float w) {
  if (p < (uint32_t)kMlpOffB1) {                       // W1[f][h]: layer-1 A operand, row h, k = f
    const uint32_t f = p / 64, h = p % 64;
    d.w1f[((h / 16) * 64 + (f / 4) * 16 + h % 16) * 4 + f % 4] = f2bf(w);
  } else if (p < (uint32_t)kMlpOffW2) {
    d.b1[p - kMlpOffB1] = w;
  } else if (p < (uint32_t)kMlpOffB2) {                // W2[i][o]
    const uint32_t q = p - kMlpOffW2, i = q / 64, o = q % 64;
    const uint16_t h = f2bf(w);
    // W2^T: row o, k-slot of unit i; W2: row i, k-slot of unit o
    auto slot = [](uint32_t unit, uint32_t& kb, uint32_t& lane_g, uint32_t& j) {
      const uint32_t blk = unit / 16, r = unit % 16;
      kb = blk / 2; lane_g = r / 4; j = (blk % 2) * 4 + r % 4;
    };
    uint32_t kb, gg, j;
    slot(i, kb, gg, j);
    d.w2tf[(((o / 16) * 2 + kb) * 64 + gg * 16 + o % 16) * 8 + j] = h;
    slot(o, kb, gg, j);
    d.w2f[(((i / 16) * 2 + kb) * 64 + gg * 16 + i % 16) * 8 + j] = h;
  } else if (p < (uint32_t)kMlpOffW3) {
    d.b2[p - kMlpOffB2] = w;
  } else if (p < (uint32_t)kMlpOffB3) {
    d.w3[p - kMlpOffW3] = w;
  } else {
    d.b3[0] = w;
  }
}

// Gradient sum over the training workgroups' partials and Adam (Kingma & Ba
// 2015).  Each workgroup takes 64 parameters (consecutive lanes: coalesced
// partial reads); its sixteen waves each sum every sixteenth partial, in
// order, four loads in flight, and the sixteen sums are added pairwise in a
// fixed tree: a fixed order, so the step is deterministic.  The last
// workgroup sums the loss partials.
constexpr int kAdamParams = 64;
constexpr int kAdamGroups = 16;
__global__ __launch_bounds__(kAdamParams * kAdamGroups) void k_mlp_adam(
    MlpState st, const float* __restrict__ part, uint32_t nwg, const float* __restrict__ loss_part, float lr,
    float beta1, float beta2, float eps, float bc1, float bc2, int apply) {
  __shared__ float red[kAdamGroups][kAdamParams];
  if (blockIdx.x == gridDim.x - 1) {
    __shared__ float lr_[kAdamParams * kAdamGroups];
    float L = 0.f;
    for (uint32_t i = threadIdx.x; i < nwg; i += blockDim.x) L += loss_part[i];
    lr_[threadIdx.x] = L;
    __syncthreads();
    for (int o = kAdamParams * kAdamGroups / 2; o >= 1; o >>= 1) {
      if ((int)threadIdx.x < o) lr_[threadIdx.x] += lr_[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) st.loss[0] = lr_[0];
    return;
  }
  const int q = threadIdx.x & (kAdamParams - 1), grp = threadIdx.x / kAdamParams;
  const uint32_t p = blockIdx.x * kAdamParams + q;
  float gsum = 0.f;
  if (p < (uint32_t)kMlpParams) {
#pragma unroll 4
    for (uint32_t i = grp; i < nwg; i += kAdamGroups) gsum += part[(size_t)i * kMlpParams + p];
  }
  red[grp][q] = gsum;
  __syncthreads();
  for (int o = kAdamGroups / 2; o >= 1; o >>= 1) {
    if (grp < o) red[grp][q] += red[grp + o][q];
    __syncthreads();
  }
  if (grp != 0 || p >= (uint32_t)kMlpParams) return;
  const float gr = red[0][q];
  st.grad[p] = gr;
  float w = st.w[p];
  if (apply) {
    float m = beta1 * st.m[p] + (1.f - beta1) * gr;
    float v = beta2 * st.v[p] + (1.f - beta2) * gr * gr;
    st.m[p] = m;
    st.v[p] = v;
    w = w - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    st.w[p] = w;
  }
  mlp_pack_one(st.dev, p, w);
}

uint32_t mlp_train_groups(uint32_t n) {
  const uint32_t chunks = (n + kTrainChunk - 1) / kTrainChunk;
  const uint32_t g = (chunks + kTrainWaves - 1) / kTrainWaves;
  return g < kMlpMaxTrainGroups ? g : kMlpMaxTrainGroups;
}

hipError_t launch_mlp_train(const MlpState& st, const float* x, const float* t, uint32_t n, float* part,
                            float* loss_part, hipStream_t s) {
  if (!x || !t || !part || !loss_part || n == 0) return hipErrorInvalidValue;
  k_mlp_train<<<mlp_train_groups(n), 64 * kTrainWaves, 0, s>>>(st.dev, x, t, n, 1.f / (float)n, part, loss_part);
  return hipGetLastError();
}
hipError_t launch_mlp_adam(const MlpState& st, const float* part, uint32_t nwg, const float* loss_part, float lr,
                           int step, int apply, hipStream_t s) {
  if (!part || !loss_part || nwg == 0) return hipErrorInvalidValue;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  const uint32_t pb = (kMlpParams + kAdamParams - 1) / kAdamParams;
  k_mlp_adam<<<pb + 1, kAdamParams * kAdamGroups, 0, s>>>(st, part, nwg, loss_part, lr, b1, b2, eps, bc1, bc2, apply);
  return hipGetLastError();
}
hipError_t launch_mlp_infer(const MlpState& st, const float* x, uint32_t n, float* y, hipStream_t s) {
  if (!x || !y) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  const uint32_t chunks = (n + 16 * kInferBlocks - 1) / (16 * kInferBlocks);
  uint32_t grid = (chunks + kTrainWaves - 1) / kTrainWaves;
  if (grid > kInferMaxGroups) grid = kInferMaxGroups;
  k_mlp_infer<<<grid, 64 * kTrainWaves, 0, s>>>(st.dev, x, n, y);
  return hipGetLastError();
}
__global__ void k_mlp_pack(MlpState st) {
  uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p < (uint32_t)kMlpParams) mlp_pack_one(st.dev, p, st.w[p]);
}
hipError_t launch_mlp_pack(const MlpState& st, hipStream_t s) {
  k_mlp_pack<<<(kMlpParams + 255) / 256, 256, 0, s>>>(st);
  return hipGetLastError();
}

// ---- renderer-side feature, target and allocation-signal kernels ----------

__device__ __forceinline__ float luminance(const float* c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

__global__ __launch_bounds__(256) void k_mlp_features(MlpFeatureArgs a, uint32_t n, float* __restrict__ x) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const size_t i3 = 3 * (size_t)i;
  const float* r = a.result + i3;
  const float* q = a.sq + i3;
  float* o = x + (size_t)i * kMlpIn;
  o[0] = log1pf(luminance(r));
  o[1] = log1pf(luminance(a.filtered + i3));
  o[2] = logf(a.err[i] + 1e-12f) * 0.1f;
  float v = ((q[0] - r[0] * r[0]) + (q[1] - r[1] * r[1]) + (q[2] - r[2] * r[2])) / 3.f;
  o[3] = logf(fmaxf(v, 0.f) + 1e-12f) * 0.1f;
  o[4] = luminance(a.gdif + i3);
  const float* qf = a.sq_f + i3;
  const float* f = a.filtered + i3;
  float vf = ((qf[0] - f[0] * f[0]) + (qf[1] - f[1] * f[1]) + (qf[2] - f[2] * f[2])) / 3.f;
  o[5] = logf(fmaxf(vf, 0.f) + 1e-12f) * 0.1f;
  float dx = a.gpos[i3] - a.cam[0], dy = a.gpos[i3 + 1] - a.cam[1], dz = a.gpos[i3 + 2] - a.cam[2];
  o[6] = log1pf(sqrtf(dx * dx + dy * dy + dz * dz)) * 0.2f;
  o[7] = logf((float)max(a.sample_n[i], 1u)) * 0.2f;
  o[8] = a.gnrm[i3 + 1];
}

__global__ __launch_bounds__(256) void k_mlp_target(const float* __restrict__ result, const float* __restrict__ ref,
                                                    const uint32_t* __restrict__ sample_n, uint32_t n, float clamp,
                                                    float* __restrict__ t) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* r = result + 3 * (size_t)i;
  const float* q = ref + 3 * (size_t)i;
  float d0 = fminf(r[0], clamp) - fminf(q[0], clamp), d1 = fminf(r[1], clamp) - fminf(q[1], clamp);
  float d2 = fminf(r[2], clamp) - fminf(q[2], clamp);
  float se = (d0 * d0 + d1 * d1 + d2 * d2) / 3.f;
  t[i] = logf((float)sample_n[i] * se + 1e-12f);
}

__global__ __launch_bounds__(256) void k_mlp_err(const float* __restrict__ y, const uint32_t* __restrict__ sample_n,
                                                 uint32_t nd, uint32_t signal, uint32_t n,
                                                 const float* __restrict__ plain, float* __restrict__ e) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t N = max(sample_n[i], 1u);
  const float v = expf(fminf(fmaxf(y[i], -60.f), 60.f));
  const float fn = (float)N;
  e[i] = signal == 3   ? sqrtf(v / fn * fmaxf(plain[i], 0.f))
         : signal == 2 ? v / fn / fn
         : signal      ? v / fn
                       : (float)nd / (float)(nd + N) * v / fn;
}

hipError_t launch_mlp_features(const MlpFeatureArgs& a, uint32_t n, float* x, hipStream_t s) {
  if (!a.result || !a.filtered || !a.err || !a.sq || !a.sq_f || !a.gpos || !a.gnrm || !a.gdif || !a.sample_n ||
      !x)
    return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  k_mlp_features<<<(n + 255) / 256, 256, 0, s>>>(a, n, x);
  return hipGetLastError();
}
hipError_t launch_mlp_target(const float* result, const float* ref, const uint32_t* sample_n, uint32_t n, float clamp,
                             float* t, hipStream_t s) {
  if (!result || !ref || !sample_n || !t) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  k_mlp_target<<<(n + 255) / 256, 256, 0, s>>>(result, ref, sample_n, n, clamp, t);
  return hipGetLastError();
}
hipError_t launch_mlp_err(const float* y, const uint32_t* sample_n, uint32_t nd, uint32_t signal, uint32_t n,
                          const float* plain, float* e, hipStream_t s) {
  if (!y || !sample_n || !e || signal > 3 || (signal == 3 && !plain)) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  k_mlp_err<<<(n + 255) / 256, 256, 0, s>>>(y, sample_n, nd, signal, n, plain, e);
  return hipGetLastError();
}

size_t mlp_lds_bytes() { return sizeof(MlpTrainLds); }

}  // namespace cpt
