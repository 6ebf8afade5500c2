// Binned-SAH BVH with NPBVHHelper semantics and its 4-wide re-layout.
//
// build_bvh_sah reproduces src/bvhhelper.cpp:55-311 bit for bit (same float
// expressions, same bin boundaries, same strict-'<' cost minimum, same Hoare
// partition with '<=' going left) but bins every triangle in one pass instead
// of rescanning the node once per bin, and emits the CompactBVH preorder
// directly.  Compiled with -ffp-contract=off so the host float arithmetic is
// the reference's.
#include <algorithm>
#include <array>
#include <limits>
#include <cmath>
#include <cstring>
#include <queue>

#include "cpt_internal.h"

namespace cpt {
namespace {

constexpr int kBins = 16;                     // BVH_SAH_BIN_N (bvhhelper.cpp:9)

struct Box {
  float mn[3], mx[3];
};
inline void box_empty(Box& b) {
  for (int k = 0; k < 3; k++) { b.mn[k] = kInfF; b.mx[k] = -kInfF; }
}
inline void box_grow(Box& b, const float* p) {
  for (int k = 0; k < 3; k++) { b.mn[k] = std::fmin(b.mn[k], p[k]); b.mx[k] = std::fmax(b.mx[k], p[k]); }
}
inline void box_merge(Box& b, const Box& o) {
  for (int k = 0; k < 3; k++) { b.mn[k] = std::fmin(b.mn[k], o.mn[k]); b.mx[k] = std::fmax(b.mx[k], o.mx[k]); }
}
// AABBSurfaceArea (bvhhelper.cpp:33-39)
inline float surface_area(const Box& b) {
  float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
  float len = std::sqrt(dx * dx + dy * dy + dz * dz);
  if (dx < 0 || dy < 0 || dz < 0 || (double)len < kEpsilon) return 0;
  return dx * dy + dx * dz + dy * dz;
}

struct Builder {
  const float* cen;        // 3 per original triangle
  const Box* tbox;         // per original triangle
  u32* order;
  std::vector<BinNode>* nodes;

  void leaf(u32 me, u32 start, u32 n) { (*nodes)[me].a = start; (*nodes)[me].n = n; }

  void build(u32 start, u32 n, u32 budget) {
    u32 me = (u32)nodes->size();
    nodes->push_back(BinNode());
    BinNode& init = (*nodes)[me];
    for (int k = 0; k < 3; k++) { init.mn[k] = kInfF; init.mx[k] = -kInfF; }
    init.a = 0; init.n = 0;
    if (n == 0) return;                                   // bvhhelper.cpp:58-59
    Box nb, cb;
    box_empty(nb); box_empty(cb);
    for (u32 i = 0; i < n; i++) {
      u32 t = order[start + i];
      box_merge(nb, tbox[t]);
      box_grow(cb, cen + (size_t)t * 3);
    }
    {
      BinNode& me_n = (*nodes)[me];
      for (int k = 0; k < 3; k++) { me_n.mn[k] = nb.mn[k]; me_n.mx[k] = nb.mx[k]; }
    }
    float dia[3] = {cb.mx[0] - cb.mn[0], cb.mx[1] - cb.mn[1], cb.mx[2] - cb.mn[2]};
    if (dia[0] < 0 || dia[1] < 0 || dia[2] < 0) return;   // :74-75
    float dlen = std::sqrt(dia[0] * dia[0] + dia[1] * dia[1] + dia[2] * dia[2]);
    if (n == 1 || budget <= 1 || (double)dlen < kEpsilon) { leaf(me, start, n); return; }
    int axis = 0;                                         // :80-83
    if (std::fabs(dia[1]) > std::fabs(dia[0])) axis = 1;
    if (std::fabs(dia[2]) > std::fabs(dia[1]) && std::fabs(dia[2]) > std::fabs(dia[0])) axis = 2;
    float len = std::fabs(dia[axis]);
    float c0 = cb.mn[axis];
    // bin boundaries exactly as :93-96; bin i = [edge[i], edge[i+1])
    float edge[kBins + 1];
    for (int i = 0; i < kBins; i++) edge[i] = c0 + (float)i * len / (float)kBins;
    edge[kBins] = (float)((double)cb.mx[axis] + kBiasEps);
    u32 binN[kBins];
    Box binB[kBins];
    for (int i = 0; i < kBins; i++) { binN[i] = 0; box_empty(binB[i]); }
    float scale = len > 0 ? (float)kBins / len : 0.f;
    for (u32 i = 0; i < n; i++) {
      u32 t = order[start + i];
      const float* c = cen + (size_t)t * 3;
      float p = c[axis];
      int k = (int)((p - c0) * scale);
      k = k < 0 ? 0 : (k > kBins - 1 ? kBins - 1 : k);
      while (k > 0 && p < edge[k]) k--;
      while (k < kBins - 1 && p >= edge[k + 1]) k++;
      if (!(p >= edge[k] && p < edge[k + 1])) continue;   // outside every bin (:101)
      binN[k]++;
      box_grow(binB[k], c);
    }
    // split costs (:120-164): heu[k] = bins 0..k left, k+1..15 right
    float nodeSA = surface_area(nb);
    u32 preN[kBins], sufN[kBins];
    Box preB[kBins], sufB[kBins];
    for (int i = 0; i < kBins; i++) {
      preN[i] = binN[i] + (i ? preN[i - 1] : 0u);
      preB[i] = binB[i];
      if (i) box_merge(preB[i], preB[i - 1]);
    }
    for (int i = kBins - 1; i >= 0; i--) {
      sufN[i] = binN[i] + (i < kBins - 1 ? sufN[i + 1] : 0u);
      sufB[i] = binB[i];
      if (i < kBins - 1) box_merge(sufB[i], sufB[i + 1]);
    }
    float minCost = kInfF, minBin = (float)(kBins - 1);
    for (int k = 0; k < kBins; k++) {
      u32 nl = preN[k], nr = k < kBins - 1 ? sufN[k + 1] : 0u;
      Box br;
      if (k < kBins - 1) br = sufB[k + 1]; else box_empty(br);
      float heu = 0.125f + ((float)nl * surface_area(preB[k]) + (float)nr * surface_area(br)) / nodeSA;
      if (heu < minCost) { minCost = heu; minBin = (float)k; }
    }
    // partition (:190-211)
    float split = c0 + (minBin + 1) * len / (float)kBins;
    u32 left = start, right = start + n - 1;
    while (left < right) {
      while (cen[(size_t)order[left] * 3 + axis] <= split && left < right) left++;
      while (cen[(size_t)order[right] * 3 + axis] > split && left < right) right--;
      if (left < right) { std::swap(order[left], order[right]); left++; right--; }
    }
    u32 mid = (cen[(size_t)order[left] * 3 + axis] <= split) ? left : left - 1;
    u32 nl = mid - start + 1;
    build(start, nl, budget - 1);
    u32 rightIdx = (u32)nodes->size();
    build(mid + 1, n - nl, budget - 1);
    (*nodes)[me].a = rightIdx - me;                       // CompactBVH offset (:294-297)
    (*nodes)[me].n = 0;
  }
};

}  // namespace

bool build_bvh_sah(const float* pos3, const u32* idx3, u32 nt, u32 max_depth, BvhHost& out) {
  out.nodes.clear();
  out.reorder.assign(nt, 0);
  if (nt == 0) return false;
  std::vector<float> cen((size_t)nt * 3);
  std::vector<Box> tb(nt);
  for (u32 i = 0; i < nt; i++) {                          // :248-259
    const float* a = pos3 + (size_t)idx3[i * 3] * 3;
    const float* b = pos3 + (size_t)idx3[i * 3 + 1] * 3;
    const float* c = pos3 + (size_t)idx3[i * 3 + 2] * 3;
    for (int k = 0; k < 3; k++) {
      cen[(size_t)i * 3 + k] = ((a[k] + b[k]) + c[k]) / 3.f;
      tb[i].mn[k] = std::fmin(c[k], std::fmin(b[k], a[k]));
      tb[i].mx[k] = std::fmax(c[k], std::fmax(b[k], a[k]));
    }
    out.reorder[i] = i;
  }
  out.nodes.reserve((size_t)nt * 2);
  Builder bld{cen.data(), tb.data(), out.reorder.data(), &out.nodes};
  bld.build(0, nt, max_depth);
  return !out.nodes.empty();
}

// ---- the device tree ----------------------------------------------------------
// The reference's SAH costs a split with the children's *centroid* bounds
// (bvhhelper.cpp:119-164), so a triangle much larger than its neighbours is
// placed by its centroid alone and every ancestor's box grows to its size: a
// ground quad spanning a scene ends up under a chain of scene-sized boxes that
// most rays must enter.  The device tree is built for traversal instead:
// 32 centroid bins on each axis costed with the triangles' real bounds, plus
// a candidate that splits off the node's oversized triangles (box surface area
// at least a quarter of the node's), down to single triangles (the 4-wide
// collapse then forms leaves); below depth 32, median splits bound the depth.
// Hits do not depend on the tree (closest = (t, original id) minimum).
namespace {

constexpr int kDevBins = 32;
constexpr u32 kDevMedianDepth = 32;

inline double half_area(const Box& b) {
  double dx = (double)b.mx[0] - b.mn[0], dy = (double)b.mx[1] - b.mn[1], dz = (double)b.mx[2] - b.mn[2];
  if (dx < 0 || dy < 0 || dz < 0) return 0.0;
  return dx * dy + dx * dz + dy * dz;
}

struct DevBuilder {
  const float* cen;
  const Box* tbox;
  u32* order;
  std::vector<BinNode>* nodes;

  void build(u32 start, u32 n, u32 depth) {
    const u32 me = (u32)nodes->size();
    nodes->push_back(BinNode());
    Box nb, cb;
    box_empty(nb);
    box_empty(cb);
    for (u32 i = 0; i < n; i++) {
      const u32 t = order[start + i];
      box_merge(nb, tbox[t]);
      box_grow(cb, cen + (size_t)t * 3);
    }
    for (int k = 0; k < 3; k++) { (*nodes)[me].mn[k] = nb.mn[k]; (*nodes)[me].mx[k] = nb.mx[k]; }
    if (n <= 1) { (*nodes)[me].a = start; (*nodes)[me].n = n; return; }
    u32* first = order + start;
    u32* last = first + n;
    u32* mid = nullptr;
    int axis = 0;
    for (int k = 1; k < 3; k++)
      if (cb.mx[k] - cb.mn[k] > cb.mx[axis] - cb.mn[axis]) axis = k;
    if (depth < kDevMedianDepth) {
      double best = INFINITY;
      int best_axis = -1, best_bin = -1;
      bool best_big = false;
      // oversized triangles against the rest
      const double big_area = 0.25 * half_area(nb);
      Box bb, rb;
      box_empty(bb);
      box_empty(rb);
      u32 nbig = 0;
      for (u32 i = 0; i < n; i++) {
        const u32 t = first[i];
        if (half_area(tbox[t]) >= big_area) { box_merge(bb, tbox[t]); nbig++; }
        else box_merge(rb, tbox[t]);
      }
      if (nbig > 0 && nbig < n) {
        best = half_area(bb) * nbig + half_area(rb) * (n - nbig);
        best_big = true;
      }
      // binned centroid splits, real bounds
      for (int ax = 0; ax < 3; ax++) {
        const float lo = cb.mn[ax], ext = cb.mx[ax] - cb.mn[ax];
        if (!(ext > 0.f)) continue;
        u32 cnt[kDevBins] = {};
        Box bx[kDevBins];
        for (int k = 0; k < kDevBins; k++) box_empty(bx[k]);
        const float scale = (float)kDevBins / ext;
        for (u32 i = 0; i < n; i++) {
          const u32 t = first[i];
          int k = (int)((cen[(size_t)t * 3 + ax] - lo) * scale);
          k = k < 0 ? 0 : (k >= kDevBins ? kDevBins - 1 : k);
          cnt[k]++;
          box_merge(bx[k], tbox[t]);
        }
        double right_area[kDevBins];
        u32 right_n[kDevBins];
        Box acc;
        box_empty(acc);
        u32 an = 0;
        for (int k = kDevBins - 1; k > 0; k--) {
          box_merge(acc, bx[k]);
          an += cnt[k];
          right_area[k] = half_area(acc);
          right_n[k] = an;
        }
        box_empty(acc);
        an = 0;
        for (int k = 0; k < kDevBins - 1; k++) {          // split after bin k
          box_merge(acc, bx[k]);
          an += cnt[k];
          if (an == 0 || right_n[k + 1] == 0) continue;
          const double c = half_area(acc) * an + right_area[k + 1] * right_n[k + 1];
          if (c < best) { best = c; best_axis = ax; best_bin = k; best_big = false; }
        }
      }
      if (best_big) {
        mid = std::partition(first, last, [&](u32 t) { return half_area(tbox[t]) >= big_area; });
      } else if (best_axis >= 0) {
        const float lo = cb.mn[best_axis], scale = (float)kDevBins / (cb.mx[best_axis] - cb.mn[best_axis]);
        mid = std::partition(first, last, [&](u32 t) {
          int k = (int)((cen[(size_t)t * 3 + best_axis] - lo) * scale);
          k = k < 0 ? 0 : (k >= kDevBins ? kDevBins - 1 : k);
          return k <= best_bin;
        });
      }
    }
    if (mid == nullptr || mid == first || mid == last) {   // median split (deep or degenerate)
      mid = first + n / 2;
      std::nth_element(first, mid, last,
                       [&](u32 x, u32 y) { return cen[(size_t)x * 3 + axis] < cen[(size_t)y * 3 + axis]; });
    }
    const u32 nl = (u32)(mid - first);
    build(start, nl, depth + 1);
    const u32 right = (u32)nodes->size();
    build(start + nl, n - nl, depth + 1);
    (*nodes)[me].a = right - me;
    (*nodes)[me].n = 0;
  }
};

}  // namespace

bool build_bvh_device(const float* pos3, const u32* idx3, u32 nt, BvhHost& out) {
  out.dnodes.clear();
  out.dreorder.resize(nt);
  if (nt == 0) return false;
  std::vector<float> cen((size_t)nt * 3);
  std::vector<Box> tb(nt);
  for (u32 i = 0; i < nt; i++) {
    const float* a = pos3 + (size_t)idx3[i * 3] * 3;
    const float* b = pos3 + (size_t)idx3[i * 3 + 1] * 3;
    const float* c = pos3 + (size_t)idx3[i * 3 + 2] * 3;
    for (int k = 0; k < 3; k++) {
      cen[(size_t)i * 3 + k] = ((a[k] + b[k]) + c[k]) / 3.f;
      tb[i].mn[k] = std::fmin(c[k], std::fmin(b[k], a[k]));
      tb[i].mx[k] = std::fmax(c[k], std::fmax(b[k], a[k]));
    }
    out.dreorder[i] = i;
  }
  out.dnodes.reserve((size_t)nt * 2);
  DevBuilder bld{cen.data(), tb.data(), out.dreorder.data(), &out.dnodes};
  bld.build(0, nt, 0);
  return true;
}

// Cost weights of the collapse below: one visit of a 4-wide node (four box
// tests, the near-first sort, push and pop) against one triangle test, in the
// ratio of their VALU instructions in the traversal loop (about 230 and 82 on
// gfx950).
constexpr double kCollapseNodeCost = 2.8, kCollapseTriCost = 1.0;

// SAH-optimal collapse of the binary device tree into 4-wide nodes, by dynamic
// programming over the binary nodes.  F[x][i] (i = 1..4) is the least expected
// cost of hanging binary subtree x below a 4-wide node as at most i entries,
// each entry a leaf range (at most leaf_max triangles, or a binary leaf):
// area(x) * tri * count, or a 4-wide node: area(x) * node + G[x][4].  G[x][i]
// spreads at most i entries over x's children (k to the left, i - k to the
// right), so F[x][i] = min(F[x][1], G[x][i]).  Children follow their parent in
// the preorder, so one reverse sweep fills both tables.  Nodes are then
// emitted breadth first from the root, whose entries are G[0][4]'s.
void collapse_bvh4(const float* /*pos3*/, const u32* /*idx3*/, BvhHost& b, u32 leaf_max) {
  const std::vector<BinNode>& bn = b.dnodes;
  u32 N = (u32)bn.size();
  std::vector<u32> cnt(N, 0), first(N, 0);
  auto is_empty = [&](u32 i) { return bn[i].n == 0 && bn[i].mn[0] > bn[i].mx[0]; };
  auto is_leaf = [&](u32 i) { return bn[i].n > 0; };
  for (u32 ii = N; ii-- > 0;) {
    if (is_empty(ii)) { cnt[ii] = 0; first[ii] = kInvalid; continue; }
    if (is_leaf(ii)) { cnt[ii] = bn[ii].n; first[ii] = bn[ii].a; continue; }
    u32 l = ii + 1, r = ii + bn[ii].a;
    cnt[ii] = cnt[l] + cnt[r];
    first[ii] = first[l] != kInvalid ? first[l] : first[r];
  }
  auto area = [&](u32 i) {
    double dx = (double)bn[i].mx[0] - bn[i].mn[0], dy = (double)bn[i].mx[1] - bn[i].mn[1];
    double dz = (double)bn[i].mx[2] - bn[i].mn[2];
    return std::max(0.0, dx * dy + dx * dz + dy * dz);
  };
  b.nodes4.clear();
  b.depth4 = 0;
  b.leaves4 = 0;
  auto set_child = [&](Node4& n4, int c, u32 binIdx) {
    const BinNode& s = bn[binIdx];
    n4.lox[c] = s.mn[0]; n4.loy[c] = s.mn[1]; n4.loz[c] = s.mn[2];
    n4.hix[c] = s.mx[0]; n4.hiy[c] = s.mx[1]; n4.hiz[c] = s.mx[2];
  };
  auto init4 = [](Node4& n4) {
    for (int c = 0; c < 4; c++) {
      n4.lox[c] = n4.loy[c] = n4.loz[c] = kInfF;
      n4.hix[c] = n4.hiy[c] = n4.hiz[c] = -kInfF;
      n4.child[c] = kInvalid;
      n4.count[c] = 0;
    }
  };
  if (N == 0) return;
  b.nodes4.push_back(Node4());
  init4(b.nodes4[0]);
  if (is_leaf(0) || cnt[0] <= leaf_max) {
    set_child(b.nodes4[0], 0, 0);
    b.nodes4[0].child[0] = first[0];
    b.nodes4[0].count[0] = cnt[0];
    b.leaves4 = 1;
    b.depth4 = 1;
    return;
  }
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<std::array<double, 5>> F(N), G(N);
  std::vector<std::array<uint8_t, 5>> gk(N);     // entries given to the left child (0: the one child takes all)
  std::vector<uint8_t> as_leaf(N, 0);
  for (u32 x = N; x-- > 0;) {
    if (cnt[x] == 0) continue;
    const double A = area(x);
    const double leafC = (is_leaf(x) || cnt[x] <= leaf_max) ? A * kCollapseTriCost * cnt[x] : inf;
    G[x].fill(inf);
    gk[x].fill(0);
    if (!is_leaf(x)) {
      const u32 l = x + 1, r = x + bn[x].a;
      for (int i = 1; i <= 4; i++) {
        if (cnt[l] && cnt[r]) {
          for (int k = 1; k < i; k++) {
            const double c = F[l][k] + F[r][i - k];
            if (c < G[x][i]) { G[x][i] = c; gk[x][i] = (uint8_t)k; }
          }
        } else {
          G[x][i] = F[cnt[l] ? l : r][i];
        }
      }
    }
    const double nodeC = is_leaf(x) ? inf : A * kCollapseNodeCost + G[x][4];
    as_leaf[x] = leafC <= nodeC;
    F[x][0] = inf;
    F[x][1] = std::min(leafC, nodeC);
    for (int i = 2; i <= 4; i++) F[x][i] = std::min(F[x][1], G[x][i]);
  }
  // the entries F[x][i] (or, for spread = true, G[x][i]) stands for
  std::vector<u32> ent;
  auto take = [&](auto&& self, u32 x, int i, bool spread) -> void {
    if (!spread && F[x][1] <= G[x][i]) { ent.push_back(x); return; }
    const u32 l = x + 1, r = x + bn[x].a;
    const int k = gk[x][i];
    if (k == 0) {
      self(self, cnt[l] ? l : r, i, false);
    } else {
      self(self, l, k, false);
      self(self, r, i - k, false);
    }
  };
  struct Item { u32 bin, slot, depth; };
  std::queue<Item> q;
  q.push({0, 0, 1});
  while (!q.empty()) {
    Item it = q.front();
    q.pop();
    b.depth4 = std::max(b.depth4, it.depth);
    ent.clear();
    take(take, it.bin, 4, true);
    Node4 n4;
    init4(n4);
    std::vector<Item> kids;
    for (int c = 0; c < (int)ent.size(); c++) {
      const u32 x = ent[c];
      set_child(n4, c, x);
      if (as_leaf[x]) {
        n4.child[c] = first[x];
        n4.count[c] = cnt[x];
        b.leaves4++;
      } else {
        u32 slot = (u32)b.nodes4.size();
        b.nodes4.push_back(Node4());
        n4.child[c] = slot;
        n4.count[c] = 0;
        kids.push_back({x, slot, it.depth + 1});
      }
    }
    b.nodes4[it.slot] = n4;
    for (auto& k : kids) q.push(k);
  }
}


// ---- 64-byte quantized nodes -------------------------------------------------

namespace {

inline float q_decode(float o, u32 q, float step) { return o + (float)q * step; }

// Smallest exponent e (step 2^(e-127)) with o + 255 * step >= hi.
u32 q_exponent(float o, float hi) {
  for (u32 e = 1; e < 255; e++) {
    float step = ldexpf(1.f, (int)e - 127);
    if (q_decode(o, 255u, step) >= hi) return e;
  }
  return 0;
}

}  // namespace

bool quantize_bvh4(BvhHost& b) {
  const std::vector<Node4>& in = b.nodes4;
  std::vector<Node4Q>& out = b.nodes4q;
  out.assign(in.size(), Node4Q());
  float qmag = 1e-30f;
  for (size_t i = 0; i < in.size(); i++) {
    const Node4& n = in[i];
    Node4Q& q = out[i];
    std::memset(&q, 0, sizeof q);
    const float* lo[3] = {n.lox, n.loy, n.loz};
    const float* hi[3] = {n.hix, n.hiy, n.hiz};
    uint8_t* qlo[3] = {q.lox, q.loy, q.loz};
    uint8_t* qhi[3] = {q.hix, q.hiy, q.hiz};
    u32 exps = 0;
    for (int a = 0; a < 3; a++) {
      float o = INFINITY, top = -INFINITY;
      for (int c = 0; c < 4; c++) {
        if (n.child[c] == kInvalid) continue;
        o = std::min(o, lo[a][c]);
        top = std::max(top, hi[a][c]);
      }
      if (!(o <= top)) { o = 0.f; top = 0.f; }          // no valid child
      u32 e = q_exponent(o, top);
      if (e == 0) return false;
      const float step = ldexpf(1.f, (int)e - 127);
      q.o[a] = o;
      qmag = std::max(qmag, std::max(std::fabs(o), std::fabs(q_decode(o, 255u, step))));
      exps |= e << (8 * a);
      for (int c = 0; c < 4; c++) {
        if (n.child[c] == kInvalid) { qlo[a][c] = 255; qhi[a][c] = 0; continue; }
        double fl = std::floor(((double)lo[a][c] - (double)o) / (double)step);
        u32 ql = (u32)std::max(0.0, std::min(255.0, fl));
        while (ql > 0 && q_decode(o, ql, step) > lo[a][c]) ql--;
        double fh = std::ceil(((double)hi[a][c] - (double)o) / (double)step);
        u32 qh = (u32)std::max(0.0, std::min(255.0, fh));
        while (qh < 255 && q_decode(o, qh, step) < hi[a][c]) qh++;
        if (q_decode(o, ql, step) > lo[a][c] || q_decode(o, qh, step) < hi[a][c]) return false;
        qlo[a][c] = (uint8_t)ql;
        qhi[a][c] = (uint8_t)qh;
      }
    }
    q.exps = exps;
    for (int c = 0; c < 4; c++) {
      if (n.count[c] > 0xFFFFu) return false;
      q.count[c] = (uint16_t)n.count[c];
      q.child[c] = n.child[c];
    }
  }
  b.qmag = qmag;
  return true;
}

}  // namespace cpt
