// Host-side BVH construction: binned-SAH binary BVH over triangle boxes, then an SAH-optimal
// collapse (dynamic programming over the binary tree) into 4-wide nodes with 8-bit quantized
// child boxes.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "lgpu_internal.h"

namespace lgpu {

// ---------------------------------------------------------------- BVH (binned SAH)
struct Box {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const float *p) { for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); } }
  void grow(const Box &b) { for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); } }
  float area() const {
    if (lo[0] > hi[0]) return 0.f;
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BvhBuilder {
  std::vector<Box> tb;
  std::vector<float> cen;  // 3 per tri
  std::vector<uint32_t> idx;
  std::vector<BvhNode> nodes;
  std::vector<std::array<Box, 2>> cbox;  // exact child boxes per node (for the 4-wide collapse)
  uint32_t maxDepth = 0;
  float pad = 0;
  static constexpr int kBins = 32;
  static constexpr uint32_t kMaxLeaf = 31;

  void storeChild(BvhNode &n, int slot, const Box &b) {
    // Conservative padding so that rounding in the slab test never culls a
    // triangle lying on a box face.
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = b.lo[a] - (std::fabs(b.lo[a]) * 4e-7f + pad);
      hi[a] = b.hi[a] + (std::fabs(b.hi[a]) * 4e-7f + pad);
    }
    if (b.lo[0] > b.hi[0]) {  // empty child: a box no ray enters
      for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
    }
    if (slot == 0) {
      n.a[0] = lo[0]; n.a[1] = hi[0]; n.a[2] = lo[1]; n.a[3] = hi[1];
      n.c[0] = lo[2]; n.c[1] = hi[2];
    } else {
      n.b[0] = lo[0]; n.b[1] = hi[0]; n.b[2] = lo[1]; n.b[3] = hi[1];
      n.c[2] = lo[2]; n.c[3] = hi[2];
    }
  }

  Box boundsOf(uint32_t first, uint32_t count) const {
    Box b;
    for (uint32_t i = first; i < first + count; ++i) b.grow(tb[idx[i]]);
    return b;
  }

  // Returns the child reference for range [first, first+count).
  uint32_t build(uint32_t first, uint32_t count, uint32_t depth) {
    maxDepth = std::max(maxDepth, depth);
    Box cb;
    for (uint32_t i = first; i < first + count; ++i) cb.grow(&cen[3 * idx[i]]);
    int bestAxis = -1, bestSplit = -1;
    float bestCost = INFINITY;
    const Box pb = boundsOf(first, count);
    const float pa = pb.area();
    if (count > 2 && depth < (uint32_t)kMaxBvhDepth - 2) {
      for (int a = 0; a < 3; ++a) {
        const float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0)) continue;
        Box bb[kBins];
        uint32_t bc[kBins] = {0};
        const float k = (float)kBins / ext;
        for (uint32_t i = first; i < first + count; ++i) {
          int b = (int)((cen[3 * idx[i] + a] - cb.lo[a]) * k);
          b = std::min(std::max(b, 0), kBins - 1);
          bb[b].grow(tb[idx[i]]);
          bc[b]++;
        }
        float rightArea[kBins];
        uint32_t rightCount[kBins];
        Box acc;
        uint32_t n = 0;
        for (int b = kBins - 1; b > 0; --b) {
          acc.grow(bb[b]);
          n += bc[b];
          rightArea[b] = acc.area();
          rightCount[b] = n;
        }
        Box la;
        uint32_t ln = 0;
        for (int b = 0; b < kBins - 1; ++b) {
          la.grow(bb[b]);
          ln += bc[b];
          if (!ln || !rightCount[b + 1]) continue;
          // SAH with a node-step cost of half a triangle test: the 4-wide collapse makes each
          // binary level cheaper to traverse than a standalone node (measured best of 0.25..3)
          const float cost = 0.5f + (la.area() * (float)ln + rightArea[b + 1] * (float)rightCount[b + 1]) / pa;
          if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = b; }
        }
      }
    }
    const bool canLeaf = count <= kMaxLeaf;
    const bool forceSplit = !canLeaf;
    if (canLeaf && (count <= 2 || bestAxis < 0 || bestCost >= (float)count || depth >= (uint32_t)kMaxBvhDepth - 2)) {
      return MakeLeaf(first, count);
    }
    uint32_t mid;
    if (bestAxis >= 0) {
      const float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
      const float k = (float)kBins / ext;
      auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](uint32_t t) {
        int b = (int)((cen[3 * t + bestAxis] - cb.lo[bestAxis]) * k);
        b = std::min(std::max(b, 0), kBins - 1);
        return b <= bestSplit;
      });
      mid = (uint32_t)(it - idx.begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    } else {
      (void)forceSplit;
      mid = first + count / 2;  // degenerate centroids: split by index
    }
    const uint32_t me = (uint32_t)nodes.size();
    nodes.push_back(BvhNode());
    cbox.emplace_back();
    const uint32_t l = build(first, mid - first, depth + 1);
    const uint32_t r = build(mid, first + count - mid, depth + 1);
    nodes[me].child[0] = l;
    nodes[me].child[1] = r;
    nodes[me].pad[0] = nodes[me].pad[1] = 0;
    cbox[me][0] = boundsOf(first, mid - first);
    cbox[me][1] = boundsOf(mid, first + count - mid);
    storeChild(nodes[me], 0, cbox[me][0]);
    storeChild(nodes[me], 1, cbox[me][1]);
    return me;
  }

  // ------------------------------------------------ 4-wide collapse with quantized boxes
  std::vector<Bvh4Node> nodes4;
  uint32_t depth4 = 0;

  void padBox(const Box &b, float lo[3], float hi[3]) const {  // as storeChild, doubled
    for (int a = 0; a < 3; ++a) {
      lo[a] = b.lo[a] - (std::fabs(b.lo[a]) * 8e-7f + 2.f * pad);
      hi[a] = b.hi[a] + (std::fabs(b.hi[a]) * 8e-7f + 2.f * pad);
    }
  }
  // Which binary nodes become 4-wide nodes: with the binary tree's leaves fixed, the SAH cost
  // of the 4-wide tree differs only by the summed area of its nodes. dpf[n][k] = least summed
  // area of the 4-wide nodes representing n's subtree in at most k child slots of its parent;
  // dps[n][k] = slots given to n's left child when n is opened into them (0: n is one slot, a
  // 4-wide node of its own). Children follow their parent in `nodes` (pre-order build).
  std::vector<std::array<float, 5>> dpf;
  std::vector<std::array<uint8_t, 5>> dps;
  float Fdp(uint32_t ref, int k) const { return (ref & kLeafBit) ? 0.f : dpf[ref][k]; }
  int bestRootSplit(uint32_t n) const {
    int ba = 1;
    float bc = INFINITY;
    for (int a = 1; a <= 3; ++a) {
      const float c = Fdp(nodes[n].child[0], a) + Fdp(nodes[n].child[1], 4 - a);
      if (c < bc) { bc = c; ba = a; }
    }
    return ba;
  }
  void solveCollapse() {
    dpf.assign(nodes.size(), {});
    dps.assign(nodes.size(), {});
    for (size_t i = nodes.size(); i-- > 0;) {
      const uint32_t l = nodes[i].child[0], r = nodes[i].child[1];
      Box b = cbox[i][0];
      b.grow(cbox[i][1]);
      const int a4 = bestRootSplit((uint32_t)i);
      dpf[i][1] = b.area() + Fdp(l, a4) + Fdp(r, 4 - a4);
      dps[i][1] = 0;
      for (int k = 2; k <= 4; ++k) {
        float bo = INFINITY;
        int bs = 0;
        for (int a = 1; a < k; ++a) {
          const float c = Fdp(l, a) + Fdp(r, k - a);
          if (c < bo) { bo = c; bs = a; }
        }
        if (bo < dpf[i][1]) { dpf[i][k] = bo; dps[i][k] = (uint8_t)bs; }
        else { dpf[i][k] = dpf[i][1]; dps[i][k] = 0; }
      }
    }
  }
  struct C { uint32_t ref; Box box; };
  void expandSlots(uint32_t ref, const Box &box, int k, C *ch, int &nc) const {
    if ((ref & kLeafBit) || k == 1 || dps[ref][k] == 0) { ch[nc++] = {ref, box}; return; }
    const int a = dps[ref][k];
    expandSlots(nodes[ref].child[0], cbox[ref][0], a, ch, nc);
    expandSlots(nodes[ref].child[1], cbox[ref][1], k - a, ch, nc);
  }
  // Child boxes of one 4-wide node: the cut of n2's binary subtree chosen by solveCollapse.
  uint32_t collapse(uint32_t n2, uint32_t depth) {
    depth4 = std::max(depth4, depth);
    if (dpf.size() != nodes.size()) solveCollapse();
    C ch[4];
    int nc = 0;
    const int a = bestRootSplit(n2);
    expandSlots(nodes[n2].child[0], cbox[n2][0], a, ch, nc);
    expandSlots(nodes[n2].child[1], cbox[n2][1], 4 - a, ch, nc);
    const uint32_t me = (uint32_t)nodes4.size();
    nodes4.emplace_back();
    uint32_t refs[4];
    for (int k = 0; k < 4; ++k) {
      if (k >= nc || ((ch[k].ref & kLeafBit) && ((ch[k].ref >> 26) & 31u) == 0) || ch[k].box.lo[0] > ch[k].box.hi[0]) {
        refs[k] = kNone;
        continue;
      }
      refs[k] = (ch[k].ref & kLeafBit) ? ch[k].ref : collapse(ch[k].ref, depth + 1);
    }
    Bvh4Node &o = nodes4[me];
    std::memset(&o, 0, sizeof(o));
    float lo[4][3], hi[4][3];
    float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 4; ++k) {
      if (refs[k] == kNone) continue;
      padBox(ch[k].box, lo[k], hi[k]);
      for (int a = 0; a < 3; ++a) { nlo[a] = std::min(nlo[a], lo[k][a]); nhi[a] = std::max(nhi[a], hi[k][a]); }
    }
    for (int a = 0; a < 3; ++a) {
      if (!(nlo[a] <= nhi[a])) { nlo[a] = 0.f; nhi[a] = 0.f; }
      // smallest power of two s with p + 255 s >= hi in float arithmetic
      const float ext = nhi[a] - nlo[a];
      int e;
      std::frexp(ext / 255.f, &e);
      float sc = std::ldexp(1.f, std::max(e, -100));
      while (nlo[a] + 255.f * sc < nhi[a]) sc *= 2.f;
      o.p[a] = nlo[a];
      o.s[a] = sc;
      uint32_t wl = 0, wh = 0;
      for (int k = 0; k < 4; ++k) {
        uint32_t ql = 255, qh = 0;  // empty slot (skipped by its kNone ref)
        if (refs[k] != kNone) {
          float fl = std::floor((lo[k][a] - nlo[a]) / sc), fh = std::ceil((hi[k][a] - nlo[a]) / sc);
          int il = (int)std::min(std::max(fl, 0.f), 255.f), ih = (int)std::min(std::max(fh, 0.f), 255.f);
          while (il > 0 && nlo[a] + (float)il * sc > lo[k][a]) --il;
          while (ih < 255 && nlo[a] + (float)ih * sc < hi[k][a]) ++ih;
          ql = (uint32_t)il;
          qh = (uint32_t)ih;
        }
        wl |= ql << (8 * k);
        wh |= qh << (8 * k);
      }
      o.q[2 * a] = wl;
      o.q[2 * a + 1] = wh;
    }
    for (int k = 0; k < 4; ++k) o.child[k] = refs[k];
    return me;
  }
};

}  // namespace lgpu
