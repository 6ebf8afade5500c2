// Per-pixel error MLP 9 -> 64 -> 64 -> 1 (GELU), BF16 operands on gfx950 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, FP32 master weights and
// Adam.  No reference counterpart: CudaPT links FANN but never calls it; the
// network replaces the filter-residual error of the adaptive renderer
// (src/cudaRTBDPTStreamAdapBilatAtomic.cu:1143-1156) as the allocation signal.
//
// One workgroup = one wave = 64 pixels.  Layer matrices are 64 x 64 and are
// computed as 4 x 4 tiles of 16 x 16; each tile is ceil(K/32) MFMAs.
//   forward   Z1 = X W1 + b1, H1 = gelu(Z1), Z2 = H1 W2 + b2, H2 = gelu(Z2), y = H2 w3 + b3
//   backward  dy = 2 (y - t) / B, dZ2 = dy w3^T * gelu'(Z2), dH1 = dZ2 W2^T,
//             dZ1 = dH1 * gelu'(Z1), dW2 = H1^T dZ2, dW1 = X^T dZ1 (all MFMA)
// Operands sit in LDS in the orientation each MFMA reads contiguously (an
// activation is stored once per orientation it is read in).  Per-wave weight
// gradients go to a partials array; k_mlp_reduce sums them in chunks of
// kReduceChunk workgroups and k_mlp_adam sums the chunk totals, both in a fixed
// order (deterministic), and applies Adam (Kingma & Ba 2015).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.h"

namespace cpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 with round-to-nearest-even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form (Hendrycks & Gimpel 2016), and its derivative.
// tanh(u) = 1 - 2 / (1 + e^(2u)) with the hardware exponential and reciprocal
// (absolute error ~1e-7, far below the BF16 rounding of the activations).
__device__ __forceinline__ float fast_tanh(float u) {
  return 1.f - __fdividef(2.f, 1.f + __expf(2.f * u));
}
__device__ __forceinline__ void gelu_fwd(float x, float& g, float& dg) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  float x2 = x * x;
  float th = fast_tanh(c * (x + a * x2 * x));
  g = 0.5f * x * (1.f + th);
  dg = 0.5f * (1.f + th) + 0.5f * x * (1.f - th * th) * c * (1.f + 3.f * a * x2);
}

// 16x16 tile of A (rows m0.., K contiguous, leading dim lda) times B^T given as
// Bt (rows n0.., K contiguous, leading dim ldb), K a multiple of 32.
// Lane l holds D[m0 + 4(l/16) + e][n0 + l%16], e = 0..3.
__device__ __forceinline__ f32x4 mfma_tile(const uint16_t* A, int lda, int m0, const uint16_t* Bt, int ldb, int n0,
                                           int K) {
  const int l = threadIdx.x & 63;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int r = l & 15, kq = (l >> 4) * 8;
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a = *(const bf16x8*)(A + (m0 + r) * lda + k0 + kq);
    bf16x8 b = *(const bf16x8*)(Bt + (n0 + r) * ldb + k0 + kq);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  return acc;
}

// LDS of the forward pass (inference): operands of the wave's 64 pixels.
struct MlpFwdLds {
  alignas(16) uint16_t w1t[64 * 32];   // W1^T  [hidden][feature], features >= 9 zero
  alignas(16) uint16_t w2t[64 * 64];   // W2^T  [out][in]
  float b1[64], b2[64], w3[64], b3;
  alignas(16) uint16_t x[64 * 32];     // X     [pixel][feature]
  alignas(16) uint16_t h1[64 * 64];    // H1    [pixel][hidden]
  float ys[64];                        // y per pixel
};
// Training adds the other operand orientations and the backward buffers.
// Training adds the other operand orientations and the backward buffers,
// within 80 KB so that two workgroups share a CU: dZ1^T reuses H1's space (H1
// is dead after the forward pass), dZ2 and dZ2^T reuse H2's (dead after dW3),
// and the GELU derivative at Z1 is recomputed in the backward pass from X.
struct MlpLds : MlpFwdLds {
  alignas(16) uint16_t w2[64 * 64];    // W2    [in][out]
  alignas(16) uint16_t xt[32 * 64];    // X^T   [feature][pixel]
  alignas(16) uint16_t h1t[64 * 64];   // H1^T  [hidden][pixel]
  union alignas(16) {
    float h2[64 * 65];                 // H2    [pixel][hidden], FP32, padded rows (forward)
    struct {
      uint16_t dz2[64 * 64];           // dZ2   [pixel][out]   (backward)
      uint16_t dz2t[64 * 64];          // dZ2^T [out][pixel]
    } bw;
  } u;
  float dy[64];
  float dg2[64 * 64];                  // GELU derivative at Z2, MFMA output layout [tile * 4 + element][lane]
};
static_assert(sizeof(MlpLds) <= 80 * 1024, "two training workgroups per CU");

template <bool TRAIN, class L>
__device__ __forceinline__ void load_weights(L& s, const MlpDevWeights& w) {
  const int l = threadIdx.x;
  for (int i = l; i < 64 * 32; i += 64) s.w1t[i] = w.w1t[i];
  for (int i = l; i < 64 * 64; i += 64) {
    s.w2t[i] = w.w2t[i];
    if constexpr (TRAIN) s.w2[i] = w.w2[i];
  }
  s.b1[l] = w.b1[l]; s.b2[l] = w.b2[l]; s.w3[l] = w.w3[l];
  if (l == 0) s.b3 = w.b3[0];
}

// Forward of the wave's 64 pixels; returns y of the lane's pixel.  With TRAIN
// the transposed operands, H2 and the GELU derivatives stay in LDS for the
// backward pass.
template <bool TRAIN, class L>
__device__ __forceinline__ float forward(L& s, const float* __restrict__ x, uint32_t base, uint32_t n) {
  const int l = threadIdx.x;
  const int col = l & 15, rq = (l >> 4) * 4;
  // X as bf16; rows past n are zero
  for (int f = 0; f < 32; f++) {
    float v = (f < kMlpIn && base + l < n) ? x[(size_t)(base + l) * kMlpIn + f] : 0.f;
    uint16_t h = f2bf(v);
    s.x[l * 32 + f] = h;
    if constexpr (TRAIN) s.xt[f * 64 + l] = h;
  }
  __syncthreads();
  // layer 1
  for (int t = 0; t < 16; t++) {
    const int m0 = (t >> 2) * 16, n0 = (t & 3) * 16;
    f32x4 z = mfma_tile(s.x, 32, m0, s.w1t, 32, n0, 32);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int px = m0 + rq + e, hd = n0 + col;
      float g, dg;
      gelu_fwd(z[e] + s.b1[hd], g, dg);
      uint16_t hb = f2bf(g);
      s.h1[px * 64 + hd] = hb;
      if constexpr (TRAIN) s.h1t[hd * 64 + px] = hb;
    }
  }
  __syncthreads();
  // layer 2, and layer 3 as a reduction over the hidden units: per row block,
  // each lane sums its 4 column tiles, then the 16 lanes of a column group add
  // up by butterfly.
  for (int mb = 0; mb < 4; mb++) {
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int t = mb * 4 + c, m0 = mb * 16, n0 = c * 16;
      f32x4 z = mfma_tile(s.h1, 64, m0, s.w2t, 64, n0, 64);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int px = m0 + rq + e, hd = n0 + col;
        float g, dg;
        gelu_fwd(z[e] + s.b2[hd], g, dg);
        part[e] = fmaf(g, s.w3[hd], part[e]);
        if constexpr (TRAIN) {
          s.dg2[(t * 4 + e) * 64 + l] = dg;
          s.u.h2[px * 65 + hd] = g;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      float v = part[e];
      for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
      if (col == e) s.ys[mb * 16 + rq + e] = v;
    }
  }
  __syncthreads();
  return s.ys[l] + s.b3;
}

__global__ __launch_bounds__(64) void k_mlp_train(MlpDevWeights w, const float* __restrict__ x,
                                                  const float* __restrict__ target, uint32_t n, float inv_batch,
                                                  float* __restrict__ part, float* __restrict__ loss_part) {
  __shared__ MlpLds s;
  const int l = threadIdx.x;
  const uint32_t base = blockIdx.x * 64;
  const int col = l & 15, rq = (l >> 4) * 4;
  load_weights<true>(s, w);
  const float y = forward<true>(s, x, base, n);
  const bool valid = base + l < n;
  const float d = valid ? (y - target[base + l]) : 0.f;
  float* P = part + (size_t)blockIdx.x * kMlpParams;
  // loss partial (sum of squares, workgroup order)
  float sq = d * d;
  for (int o = 32; o >= 1; o >>= 1) sq += __shfl_xor(sq, o, 64);
  if (l == 0) loss_part[blockIdx.x] = sq;
  s.dy[l] = 2.f * d * inv_batch;
  __syncthreads();
  // dW3 (lane = hidden unit), db3
  {
    float g3 = 0.f;
    for (int px = 0; px < 64; px++) g3 = fmaf(s.u.h2[px * 65 + l], s.dy[px], g3);
    P[kMlpOffW3 + l] = g3;
    float db3 = s.dy[l];
    for (int o = 32; o >= 1; o >>= 1) db3 += __shfl_xor(db3, o, 64);
    if (l == 0) P[kMlpOffB3] = db3;
  }
  __syncthreads();                                 // H2 is dead: its space takes dZ2
  // dZ2 = dy w3^T * gelu'(Z2), both orientations; db2 = column sums
  float db2[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
  for (int t = 0; t < 16; t++) {
    const int m0 = (t >> 2) * 16, n0 = (t & 3) * 16;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int px = m0 + rq + e, hd = n0 + col;
      float g = s.dy[px] * s.w3[hd] * s.dg2[(t * 4 + e) * 64 + l];
      db2[t & 3] += g;
      uint16_t hb = f2bf(g);
      s.u.bw.dz2[px * 64 + hd] = hb;
      s.u.bw.dz2t[hd * 64 + px] = hb;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    float v = db2[c];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (l < 16) P[kMlpOffB2 + c * 16 + l] = v;
  }
  __syncthreads();
  // dH1 = dZ2 W2^T -> dZ1 = dH1 * gelu'(Z1) (transposed, into H1's space), db1;
  // Z1 is recomputed by the same MFMA as in the forward pass
  uint16_t* dz1t = s.h1;
  float db1[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
  for (int t = 0; t < 16; t++) {
    const int m0 = (t >> 2) * 16, n0 = (t & 3) * 16;
    f32x4 dh = mfma_tile(s.u.bw.dz2, 64, m0, s.w2, 64, n0, 64);
    f32x4 z = mfma_tile(s.x, 32, m0, s.w1t, 32, n0, 32);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int px = m0 + rq + e, hd = n0 + col;
      float g1, dg1;
      gelu_fwd(z[e] + s.b1[hd], g1, dg1);
      float g = dh[e] * dg1;
      db1[t & 3] += g;
      dz1t[hd * 64 + px] = f2bf(g);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    float v = db1[c];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (l < 16) P[kMlpOffB1 + c * 16 + l] = v;
  }
  __syncthreads();
  // dW2[in][out] = H1^T dZ2 : A = H1^T [in][px], Bt = dZ2^T [out][px]
  for (int t = 0; t < 16; t++) {
    const int m0 = (t >> 2) * 16, n0 = (t & 3) * 16;
    f32x4 g = mfma_tile(s.h1t, 64, m0, s.u.bw.dz2t, 64, n0, 64);
#pragma unroll
    for (int e = 0; e < 4; e++) P[kMlpOffW2 + (m0 + rq + e) * 64 + n0 + col] = g[e];
  }
  // dW1[feature][hidden] = X^T dZ1 : A = X^T [feature][px] (rows 0..15), Bt = dZ1^T [hidden][px]
  for (int c = 0; c < 4; c++) {
    f32x4 g = mfma_tile(s.xt, 64, 0, dz1t, 64, c * 16, 64);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int f = rq + e;
      if (f < kMlpIn) P[kMlpOffW1 + f * 64 + c * 16 + col] = g[e];
    }
  }
}

__global__ __launch_bounds__(64) void k_mlp_infer(MlpDevWeights w, const float* __restrict__ x, uint32_t n,
                                                  float* __restrict__ y) {
  __shared__ MlpFwdLds s;
  const uint32_t base = blockIdx.x * 64;
  load_weights<false>(s, w);
  const float out = forward<false>(s, x, base, n);
  if (base + threadIdx.x < n) y[base + threadIdx.x] = out;
}

// Writes parameter p (value w) into the BF16/FP32 operand pack.
__device__ __forceinline__ void mlp_pack_one(MlpDevWeights d, uint32_t p, float w) {
  if (p < (uint32_t)kMlpOffB1) {                       // W1[f][h] -> W1^T[h][f]
    uint32_t f = p / 64, h = p % 64;
    d.w1t[h * 32 + f] = f2bf(w);
  } else if (p < (uint32_t)kMlpOffW2) {
    d.b1[p - kMlpOffB1] = w;
  } else if (p < (uint32_t)kMlpOffB2) {                // W2[i][o]
    uint32_t q = p - kMlpOffW2, i = q / 64, o = q % 64;
    uint16_t h = f2bf(w);
    d.w2[i * 64 + o] = h;
    d.w2t[o * 64 + i] = h;
  } else if (p < (uint32_t)kMlpOffW3) {
    d.b2[p - kMlpOffB2] = w;
  } else if (p < (uint32_t)kMlpOffB3) {
    d.w3[p - kMlpOffW3] = w;
  } else {
    d.b3[0] = w;
  }
}

// Sum of the per-workgroup gradients (workgroup order), Adam, and the BF16
// operand copies in the layouts the MFMA tiles read.
// Chunk c of the per-workgroup gradients: part2[c][p] = sum over workgroups
// c*kReduceChunk .. of part[wg][p], in workgroup order.
constexpr uint32_t kReduceChunk = 16;
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, uint32_t nwg,
                                                    float* __restrict__ part2) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= (uint32_t)kMlpParams) return;
  const uint32_t i0 = blockIdx.y * kReduceChunk, i1 = min(i0 + kReduceChunk, nwg);
  float g = 0.f;
  for (uint32_t i = i0; i < i1; i++) g += part[(size_t)i * kMlpParams + p];
  part2[(size_t)blockIdx.y * kMlpParams + p] = g;
}

// The last workgroup sums the loss partials (strided per thread, then a fixed
// tree); the others sum the chunk totals of one parameter each and step Adam.
__global__ __launch_bounds__(256) void k_mlp_adam(MlpState st, const float* __restrict__ part2, uint32_t nch,
                                                  const float* __restrict__ loss_part, uint32_t nwg, float lr,
                                                  float beta1, float beta2, float eps, float bc1, float bc2,
                                                  int apply) {
  if (blockIdx.x == gridDim.x - 1) {
    __shared__ float red[256];
    float L = 0.f;
    for (uint32_t i = threadIdx.x; i < nwg; i += 256) L += loss_part[i];
    red[threadIdx.x] = L;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) st.loss[0] = red[0];
    return;
  }
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= (uint32_t)kMlpParams) return;
  float g = 0.f;
  for (uint32_t c = 0; c < nch; c++) g += part2[(size_t)c * kMlpParams + p];
  st.grad[p] = g;
  float w = st.w[p];
  if (apply) {
    float m = beta1 * st.m[p] + (1.f - beta1) * g;
    float v = beta2 * st.v[p] + (1.f - beta2) * g * g;
    st.m[p] = m;
    st.v[p] = v;
    w = w - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    st.w[p] = w;
  }
  mlp_pack_one(st.dev, p, w);
}

hipError_t launch_mlp_train(const MlpState& st, const float* x, const float* t, uint32_t n, float* part,
                            float* loss_part, hipStream_t s) {
  if (!x || !t || !part || !loss_part || n == 0) return hipErrorInvalidValue;
  uint32_t nwg = (n + 63) / 64;
  k_mlp_train<<<nwg, 64, 0, s>>>(st.dev, x, t, n, 1.f / (float)n, part, loss_part);
  return hipGetLastError();
}
hipError_t launch_mlp_adam(const MlpState& st, const float* part, uint32_t nwg, const float* loss_part, float lr,
                           int step, int apply, float* part2, hipStream_t s) {
  if (!part || !loss_part || !part2 || nwg == 0) return hipErrorInvalidValue;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  const uint32_t nch = mlp_reduce_chunks(nwg);
  const uint32_t pb = (kMlpParams + 255) / 256;
  k_mlp_reduce<<<dim3(pb, nch), 256, 0, s>>>(part, nwg, part2);
  k_mlp_adam<<<pb + 1, 256, 0, s>>>(st, part2, nch, loss_part, nwg, lr, b1, b2, eps, bc1, bc2, apply);
  return hipGetLastError();
}
hipError_t launch_mlp_infer(const MlpState& st, const float* x, uint32_t n, float* y, hipStream_t s) {
  if (!x || !y) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  k_mlp_infer<<<(n + 63) / 64, 64, 0, s>>>(st.dev, x, n, y);
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

size_t mlp_lds_bytes() { return sizeof(MlpLds); }
uint32_t mlp_reduce_chunks(uint32_t nwg) { return (nwg + kReduceChunk - 1) / kReduceChunk; }

}  // namespace cpt
