// Internal layouts shared by the host library and the gfx950 kernels.
#pragma once
#include <cstdint>
#include <vector>

typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;

namespace cpt {

// Reference constants (include/cudahelper.h:27-30).
constexpr float kInfF = 1e20f;                 // M_INF as float
constexpr double kBiasEps = 1e-4;              // M_FLT_BIAS_EPSILON
constexpr double kFltEps = 1e-5;               // M_FLT_EPSILON
constexpr double kEpsilon = 1e-9;              // M_EPSILON
constexpr u32 kInvalid = 0xFFFFFFFFu;

// Compact binary node of the reference layout (src/cudaRTCommon.cu:527-541):
// min/max bounds + (offset-or-triStart, triN).  triN == 0 marks an interior
// node; an empty leaf (a degenerate split) has inverted bounds.
struct BinNode {
  float mn[3], mx[3];
  u32 a, n;
};

// 4-wide node, 128 bytes = one 128-byte line: six 4-float slab planes, then
// per child either an interior node index (count == 0) or a leaf triangle
// range (child = first leaf-order triangle, count = #triangles).
// Unused slots have child = kInvalid and an empty box.
struct alignas(128) Node4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  u32 child[4];
  u32 count[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

// The same node in 64 bytes: child boxes quantized to 8 bits on a per-node
// grid (origin o, power-of-two step 2^(e-127) per axis), decoded as
// o + q * step in FP32.  The host picks every q so that the decoded box
// contains the child's FP32 box (lo rounded down, hi up, checked with the
// decode arithmetic), so culling stays conservative and hits unchanged.
// Loaded as four 16-byte vectors: {o, exponents}, {lo/hi x, lo/hi y},
// {lo/hi z, counts}, {children}.  Counts above 65535 cannot be stored (the
// build reports it).
struct alignas(64) Node4Q {
  float o[3];
  u32 exps;            // ex | ey << 8 | ez << 16
  uint8_t lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  uint16_t count[4];
  u32 child[4];
};
static_assert(sizeof(Node4Q) == 64, "Node4Q must be 64 bytes");

// Host result of the SAH build.
struct BvhHost {
  std::vector<BinNode> nodes;      // DFS preorder, reference CompactBVH order
  std::vector<u32> reorder;        // leaf order -> original triangle id
  std::vector<BinNode> dnodes;     // the device tree's binary SAH build, same format
  std::vector<u32> dreorder;       // device leaf order -> original triangle id
  std::vector<Node4> nodes4;       // 4-wide re-layout (BFS order, root 0)
  std::vector<Node4Q> nodes4q;     // the same nodes quantized to 64 bytes
  float qmag = 1e-30f;             // max |o + q * step| over nodes4q's planes, at least 1e-30
  u32 depth4 = 0, leaves4 = 0;
};

// SAH build with NPBVHHelper semantics (src/bvhhelper.cpp:55-311).
bool build_bvh_sah(const float* pos3, const u32* idx3, u32 nt, u32 max_depth, BvhHost& out);
// The device tree: binned SAH over triangle bounds (dnodes / dreorder).
bool build_bvh_device(const float* pos3, const u32* idx3, u32 nt, BvhHost& out);
// Collapse the device binary tree into 4-wide nodes (SAH-optimal); a leaf
// range holds at most leaf_max triangles.
void collapse_bvh4(const float* pos3, const u32* idx3, BvhHost& b, u32 leaf_max);
// Quantize nodes4 into nodes4q; false when a node cannot be encoded.
bool quantize_bvh4(BvhHost& b);

}  // namespace cpt
